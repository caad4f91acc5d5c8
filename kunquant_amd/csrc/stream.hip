// Stream contexts: HBM-resident per-tick execution of a STREAM module
// (replaces StreamContext, cpp/Kun/Runtime.cpp:424-546).
//
// Every module buffer is one device row of num_stocks values (the newest
// row); each time-series stage owns its carried state -- history rings of
// the values it reads through windows and the state of its stateful ops --
// as [slot][S] arrays in HBM, initialised by the stage's generated init
// kernel. run() enqueues one tick kernel per stage (Rank/Scale: the row
// kernels on a single row) in stage-DAG order.
//
// Serialised state: the reference's byte layout (Runtime.cpp:534-546), built
// from the module's wire table -- every buffer of the reference's buffer
// table as a StreamBuffer (rows of history, then the per-block position),
// then every state buffer. Rows come from the module buffer (newest row) or
// the stage ring holding that value's history, state fields from the
// stage's state columns. A blob is taken between ticks, after every input
// of the last tick was pushed. Short input while reading the buffers raises
// "Failed to read initial stream buffer", short input in the state section
// "Failed to deserialize state buffer" (Runtime.cpp:404-406, 454-458).
#include <hip/hip_runtime.h>
#include <string.h>

#include <algorithm>
#include <fstream>
#include <iterator>
#include <limits>
#include <string>
#include <vector>

#include "kun_cs.hpp"
#include "kun_device.hpp"
#include "kun_module_abi.h"
#include "kun_rt.hpp"

namespace kunrt {

template <typename T> __global__ void fill_nan(T *p, long n) {
    long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = kun::qnan<T>();
}

struct StreamCtx {
    std::shared_ptr<Executor> ex;
    const KunModuleDesc *m = nullptr;
    size_t S = 0, esz = 4, tick = 0;
    std::vector<void *> rows;       // device rows per buffer
    std::vector<void *> host_rows;  // pinned mirrors for getCurrentBuffer
    std::vector<char> host_valid;
    std::vector<void *> state;      // per stage: [state_bytes * S]
    std::vector<size_t> state_len;
    // wire positions: buffer j of the wire table is at reference position
    // pos0[j] + (tick - tick0) (mod its window); set when resuming
    size_t tick0 = 0;
    std::vector<uint64_t> pos0;
    ~StreamCtx() {
        if (ex) hipStreamSynchronize(ex->stream);
        for (void *p : rows) if (p) hipFree(p);
        for (void *p : host_rows) if (p) hipHostFree(p);
        for (void *p : state) if (p) hipFree(p);
    }
};

static void launch_stage_kernel(const void *fn, StreamCtx *c, int si, long tick) {
    const KunStageDesc &sd = c->m->stages[si];
    kun::StreamArgs a;
    memset(&a, 0, sizeof a);
    int n = 0;
    for (int j = 0; j < sd.nin; j++) a.bufs[n++] = c->rows[sd.in[j]];
    for (int j = 0; j < sd.nout; j++) a.bufs[n++] = c->rows[sd.out[j]];
    size_t tb = c->state_len[si];
    (void)tb;
    a.st = c->state[si];
    // integer slots follow the T slots; the stage's byte count per stock is
    // nT*esz + nI*4, and the emitter lays T slots first
    a.ist = nullptr;
    a.S = (long)c->S;
    a.tick = tick;
    void *args[] = {&a};
    unsigned grid = (unsigned)((c->S + sd.block - 1) / sd.block);
    HIPCHECK(hipLaunchKernel(fn, dim3(grid), dim3(sd.block), args, 0, c->ex->stream));
}

// ---- reference wire format ---------------------------------------------
static void pull_state(StreamCtx *c, WireHost &h) {
    HIPCHECK(hipStreamSynchronize(c->ex->stream));
    h.rows.assign(c->m->nbuf, std::vector<char>(c->esz * c->S));
    for (int b = 0; b < c->m->nbuf; b++)
        HIPCHECK(hipMemcpy(h.rows[b].data(), c->rows[b], h.rows[b].size(), hipMemcpyDeviceToHost));
    h.state.assign(c->m->nstage, std::vector<char>());
    for (int si = 0; si < c->m->nstage; si++) {
        h.state[si].resize(c->state_len[si]);
        if (c->state_len[si])
            HIPCHECK(hipMemcpy(h.state[si].data(), c->state[si], c->state_len[si], hipMemcpyDeviceToHost));
    }
}

static void push_state(StreamCtx *c, const WireHost &h) {
    for (int b = 0; b < c->m->nbuf; b++)
        HIPCHECK(hipMemcpy(c->rows[b], h.rows[b].data(), h.rows[b].size(), hipMemcpyHostToDevice));
    for (int si = 0; si < c->m->nstage; si++)
        if (c->state_len[si])
            HIPCHECK(hipMemcpy(c->state[si], h.state[si].data(), c->state_len[si], hipMemcpyHostToDevice));
}

static inline long wmod(long x, long m) {
    long r = x % m;
    return r < 0 ? r + m : r;
}

// T slot `slot` of a stage state ([slot][S] T columns, then int32 columns)
template <typename T> static T *tcol(std::vector<char> &st, long slot, size_t S) {
    return (T *)st.data() + (size_t)slot * S;
}
template <typename T> static int32_t *icol(std::vector<char> &st, int tslots, long slot, size_t S) {
    return (int32_t *)((T *)st.data() + (size_t)tslots * S) + (size_t)slot * S;
}

static const KunRingRef *longest_ring(const KunWireBuf &wb) {
    const KunRingRef *r = nullptr;
    for (int i = 0; i < wb.nrings; i++)
        if (!r || wb.rings[i].slots > r->slots) r = &wb.rings[i];
    return r;
}

// IT: the integer type of the reference's per-lane NaN counters (int32 for
// float modules, int64 for double)
template <typename T, typename IT> static void wire_write(const WireView &c, WireHost &h, std::string &out) {
    const KunModuleDesc *m = c.m;
    const size_t S = c.S, B = (size_t)m->blocking, nblk = (S + B - 1) / B;
    const long n = c.tick;
    std::vector<T> data;
    std::vector<uint64_t> posv(nblk);
    for (int j = 0; j < m->nwire_buf; j++) {
        const KunWireBuf &wb = m->wire_bufs[j];
        const long W = wb.window;
        const long pos = W == 1 ? 0 : wmod((long)(*c.pos0)[j] + n - c.tick0, W);
        const KunRingRef *r = longest_ring(wb);
        data.resize((size_t)W * S);
        for (long p = 0; p < W; p++) {
            const long a = wmod(pos - 1 - p, W);  // age of row p
            T *dst = data.data() + p * S;
            if (wb.buf >= 0 && a == 0 && (W == 1 || !r)) {
                memcpy(dst, h.rows[wb.buf].data(), S * sizeof(T));
            } else {
                if (!r) throw Error(std::string("no history kept for stream buffer ") + wb.name);
                const long q = (n - 1 - a) & (long)(r->slots - 1);
                memcpy(dst, tcol<T>(h.state[r->stage], r->tslot + q, S), S * sizeof(T));
            }
        }
        out.append((const char *)data.data(), data.size() * sizeof(T));
        std::fill(posv.begin(), posv.end(), (uint64_t)pos);
        out.append((const char *)posv.data(), nblk * sizeof(uint64_t));
    }
    for (int j = 0; j < m->nwire_state; j++) {
        const KunWireState &ws = m->wire_states[j];
        std::vector<char> &st = h.state[ws.stage];
        auto putT = [&](int slot, size_t s0) { out.append((const char *)(tcol<T>(st, slot, S) + s0), B * sizeof(T)); };
        auto putI = [&](int slot, size_t s0) {
            const int32_t *col = icol<T>(st, ws.tslots, slot, S);
            for (size_t l = 0; l < B; l++) {
                IT v = (IT)col[s0 + l];
                out.append((const char *)&v, sizeof v);
            }
        };
        for (size_t blk = 0; blk < nblk; blk++) {
            const size_t s0 = blk * B;
            switch (ws.kind) {
            case KUN_WS_FASTSUM:  // v, compensationAdd, compensationSub, num_nans (Ops.hpp:331-334)
                for (int f = 0; f < 3; f++) putT(ws.tslot + f, s0);
                putI(ws.islot, s0);
                break;
            case KUN_WS_LINREG:  // i, x, x2, y, y2, xy sums, num_nans (Ops.hpp:382-388)
                for (int f = 0; f < 6; f++) putT(ws.tslot + f, s0);
                putI(ws.islot, s0);
                break;
            case KUN_WS_EMA:
            case KUN_WS_ACCUM:
                putT(ws.tslot, s0);
                break;
            case KUN_WS_SKIPLIST: {  // SkipList.cpp:299-352
                const int W = ws.window;
                const int32_t *cnt = icol<T>(st, ws.tslots, ws.islot + W + 1, S);
                const int32_t *last = icol<T>(st, ws.tslots, ws.islot + W + 2, S);
                out.append((const char *)(last + s0), B * sizeof(int32_t));
                for (size_t l = 0; l < B; l++) {
                    const int32_t k = cnt[s0 + l];
                    out.append((const char *)&k, sizeof k);
                    for (int q = 0; q < k; q++) {
                        const double v = (double)tcol<T>(st, ws.tslot + q, S)[s0 + l];
                        const uint64_t ix = (uint64_t)(int64_t)icol<T>(st, ws.tslots, ws.islot + q, S)[s0 + l];
                        out.append((const char *)&v, sizeof v);
                        out.append((const char *)&ix, sizeof ix);
                    }
                }
                break;
            }
            default:
                throw Error("unknown stream state kind");
            }
        }
    }
}

// Restores rows, rings and op state from a blob in the reference's layout;
// `h` holds the freshly initialised state. The owner restarts its tick
// count at 0: age a of a value goes to ring slot (-1 - a) & (slots - 1).
template <typename T, typename IT>
static void wire_read(const WireView &c, WireHost &h, const char *p, const char *end) {
    const KunModuleDesc *m = c.m;
    const size_t S = c.S, B = (size_t)m->blocking, nblk = (S + B - 1) / B;
    auto take = [&](void *dst, size_t n) {
        if ((size_t)(end - p) < n) return false;
        memcpy(dst, p, n);
        p += n;
        return true;
    };
    const T nan = std::numeric_limits<T>::quiet_NaN();
    std::vector<T> data;
    std::vector<uint64_t> posv(nblk);
    for (int j = 0; j < m->nwire_buf; j++) {
        const KunWireBuf &wb = m->wire_bufs[j];
        const long W = wb.window;
        data.resize((size_t)W * S);
        if (!take(data.data(), data.size() * sizeof(T)) || !take(posv.data(), nblk * sizeof(uint64_t)))
            throw Error("Failed to read initial stream buffer");
        for (size_t blk = 0; blk < nblk; blk++)
            if (posv[blk] != posv[0] || posv[blk] >= (uint64_t)W) throw Error("Failed to read initial stream buffer");
        const long pos = (long)posv[0];
        (*c.pos0)[j] = (uint64_t)pos;
        auto age = [&](long a) { return data.data() + wmod(pos - 1 - a, W) * S; };
        if (wb.buf >= 0) memcpy(h.rows[wb.buf].data(), age(0), S * sizeof(T));
        for (int i = 0; i < wb.nrings; i++) {
            const KunRingRef &r = wb.rings[i];
            for (long a = 0; a < r.slots; a++) {
                T *dst = tcol<T>(h.state[r.stage], r.tslot + ((-1 - a) & (long)(r.slots - 1)), S);
                if (a < W)
                    memcpy(dst, age(a), S * sizeof(T));
                else
                    std::fill(dst, dst + S, nan);
            }
        }
    }
    for (int j = 0; j < m->nwire_state; j++) {
        const KunWireState &ws = m->wire_states[j];
        std::vector<char> &st = h.state[ws.stage];
        auto getT = [&](int slot, size_t s0) {
            if (!take(tcol<T>(st, slot, S) + s0, B * sizeof(T))) throw Error("Failed to deserialize state buffer");
        };
        auto getI = [&](int slot, size_t s0) {
            int32_t *col = icol<T>(st, ws.tslots, slot, S);
            for (size_t l = 0; l < B; l++) {
                IT v;
                if (!take(&v, sizeof v)) throw Error("Failed to deserialize state buffer");
                col[s0 + l] = (int32_t)v;
            }
        };
        for (size_t blk = 0; blk < nblk; blk++) {
            const size_t s0 = blk * B;
            switch (ws.kind) {
            case KUN_WS_FASTSUM:
                for (int f = 0; f < 3; f++) getT(ws.tslot + f, s0);
                getI(ws.islot, s0);
                break;
            case KUN_WS_LINREG:
                for (int f = 0; f < 6; f++) getT(ws.tslot + f, s0);
                getI(ws.islot, s0);
                break;
            case KUN_WS_EMA:
            case KUN_WS_ACCUM:
                getT(ws.tslot, s0);
                break;
            case KUN_WS_SKIPLIST: {
                const int W = ws.window;
                int32_t *cnt = icol<T>(st, ws.tslots, ws.islot + W + 1, S);
                int32_t *last = icol<T>(st, ws.tslots, ws.islot + W + 2, S);
                if (!take(last + s0, B * sizeof(int32_t))) throw Error("Failed to deserialize state buffer");
                for (size_t l = 0; l < B; l++) {
                    int32_t k;
                    if (!take(&k, sizeof k) || k < 0 || k > W) throw Error("Failed to deserialize state buffer");
                    cnt[s0 + l] = k;
                    for (int q = 0; q < k; q++) {
                        double v;
                        uint64_t ix;
                        if (!take(&v, sizeof v) || !take(&ix, sizeof ix))
                            throw Error("Failed to deserialize state buffer");
                        tcol<T>(st, ws.tslot + q, S)[s0 + l] = (T)v;
                        icol<T>(st, ws.tslots, ws.islot + q, S)[s0 + l] = (int32_t)ix;
                    }
                }
                break;
            }
            default:
                throw Error("unknown stream state kind");
            }
        }
    }
}

void wire_serialize(const WireView &v, WireHost &h, std::string &out) {
    if (elem_size(v.m) == 8)
        wire_write<double, int64_t>(v, h, out);
    else
        wire_write<float, int32_t>(v, h, out);
}

void wire_deserialize(const WireView &v, WireHost &h, const char *p, const char *end) {
    if (elem_size(v.m) == 8)
        wire_read<double, int64_t>(v, h, p, end);
    else
        wire_read<float, int32_t>(v, h, p, end);
}

StreamCtx *stream_create(std::shared_ptr<Executor> ex, const KunModuleDesc *m, size_t S,
                         const char *state_src, size_t state_len, bool from_file) {
    if (!m || m->abi != KUN_MODULE_ABI_VERSION)
        throw Error("The required version in the module does not match the runtime version");
    if (m->out_layout != KUN_LAYOUT_STREAM) throw Error("Cannot run batch mode module via StreamContext");
    if (m->aligned && S % (size_t)m->blocking) throw Error("num_stocks must be a multiple of blocking_len");
    std::unique_ptr<StreamCtx> c(new StreamCtx());
    c->ex = ex;
    c->m = m;
    c->S = S;
    c->esz = elem_size(m);
    std::lock_guard<std::mutex> g(ex->mu);
    HIPCHECK(hipSetDevice(ex->device));
    c->rows.resize(m->nbuf, nullptr);
    c->host_rows.resize(m->nbuf, nullptr);
    c->host_valid.assign(m->nbuf, 0);
    for (int b = 0; b < m->nbuf; b++) {
        HIPCHECK(hipMalloc(&c->rows[b], c->esz * S));
        HIPCHECK(hipHostMalloc(&c->host_rows[b], c->esz * S, hipHostMallocDefault));
        unsigned grid = (unsigned)((S + 255) / 256);
        if (c->esz == 8)
            hipLaunchKernelGGL(fill_nan<double>, dim3(grid), dim3(256), 0, ex->stream, (double *)c->rows[b], (long)S);
        else
            hipLaunchKernelGGL(fill_nan<float>, dim3(grid), dim3(256), 0, ex->stream, (float *)c->rows[b], (long)S);
        HIPCHECK(hipGetLastError());
    }
    c->state.resize(m->nstage, nullptr);
    c->state_len.resize(m->nstage, 0);
    for (int si = 0; si < m->nstage; si++) {
        const KunStageDesc &sd = m->stages[si];
        if (sd.kind != KUN_STAGE_TS || !sd.state_bytes) continue;
        c->state_len[si] = (size_t)sd.state_bytes * S;
        HIPCHECK(hipMalloc(&c->state[si], c->state_len[si]));
        launch_stage_kernel(m->stream_inits[si], c.get(), si, 0);
    }
    HIPCHECK(hipStreamSynchronize(ex->stream));
    c->pos0.assign(m->nwire_buf, 0);
    if (state_src) {
        std::string blob;
        if (from_file) {
            std::ifstream f(state_src, std::ios::binary);
            blob.assign(std::istreambuf_iterator<char>(f), std::istreambuf_iterator<char>());
        } else {
            blob.assign(state_src, state_len);
        }
        WireHost h;
        pull_state(c.get(), h);
        wire_deserialize(WireView{m, S, 0, 0, &c->pos0}, h, blob.data(), blob.data() + blob.size());
        push_state(c.get(), h);
        c->tick = 0;
        c->tick0 = 0;
    }
    return c.release();
}

void stream_destroy(StreamCtx *c) { delete c; }

size_t stream_query(StreamCtx *c, const char *name) {
    for (int b = 0; b < c->m->nbuf; b++)
        if (!strcmp(c->m->bufs[b].name, name)) return (size_t)b;
    throw Error("Cannot find the buffer name");
}

void stream_push(StreamCtx *c, size_t handle, const void *host_row) {
    if (handle >= (size_t)c->m->nbuf) throw Error("bad buffer handle");
    std::lock_guard<std::mutex> g(c->ex->mu);
    HIPCHECK(hipMemcpyAsync(c->rows[handle], host_row, c->esz * c->S, hipMemcpyHostToDevice, c->ex->stream));
    HIPCHECK(hipStreamSynchronize(c->ex->stream));
    c->host_valid[handle] = 0;
}

void stream_run(StreamCtx *c) {
    const KunModuleDesc *m = c->m;
    std::lock_guard<std::mutex> g(c->ex->mu);
    HIPCHECK(hipSetDevice(c->ex->device));
    for (int si : stage_order(m)) {
        const KunStageDesc &sd = m->stages[si];
        if (sd.kind == KUN_STAGE_TS) {
            launch_stage_kernel(sd.stream_kernel, c, si, (long)c->tick);
        } else if (sd.kind == KUN_STAGE_GENERIC) {
            kun::CsArgs a;
            memset(&a, 0, sizeof a);
            int n = 0;
            for (int j = 0; j < sd.nin; j++) a.bufs[n++] = c->rows[sd.in[j]];
            for (int j = 0; j < sd.nout; j++) a.bufs[n++] = c->rows[sd.out[j]];
            for (int j = 0; j < n; j++) {
                a.lay[j] = 2;
                a.rows[j] = 1;
            }
            a.S = (long)c->S;
            a.L = 1;
            a.blk = m->blocking;
            void *args[] = {&a};
            HIPCHECK(hipLaunchKernel(sd.stream_kernel, dim3(1), dim3(64), args, 0, c->ex->stream));
        } else {
            kun_cs::View vi{c->rows[sd.in[0]], 2, 1, 0}, vo{c->rows[sd.out[0]], 2, 1, 0};
            hipError_t e = sd.kind == KUN_STAGE_RANK
                               ? kun_cs::launch_rank(m->dtype, vi, vo, (long)c->S, 1, m->blocking, c->ex->stream)
                               : kun_cs::launch_scale(m->dtype, vi, vo, (long)c->S, 1, m->blocking, c->ex->stream);
            if (e != hipSuccess) throw Error(std::string("cross-sectional kernel: ") + hipGetErrorString(e));
        }
    }
    HIPCHECK(hipStreamSynchronize(c->ex->stream));
    std::fill(c->host_valid.begin(), c->host_valid.end(), 0);
    c->tick++;
}

const void *stream_current(StreamCtx *c, size_t handle) {
    if (handle >= (size_t)c->m->nbuf) throw Error("bad buffer handle");
    if (!c->host_valid[handle]) {
        std::lock_guard<std::mutex> g(c->ex->mu);
        HIPCHECK(hipMemcpyAsync(c->host_rows[handle], c->rows[handle], c->esz * c->S, hipMemcpyDeviceToHost,
                                c->ex->stream));
        HIPCHECK(hipStreamSynchronize(c->ex->stream));
        c->host_valid[handle] = 1;
    }
    return c->host_rows[handle];
}

const void *stream_current_device(StreamCtx *c, size_t handle) {
    if (handle >= (size_t)c->m->nbuf) throw Error("bad buffer handle");
    return c->rows[handle];
}

size_t stream_ticks(StreamCtx *c) { return c->tick; }

std::string stream_serialize(StreamCtx *c) {
    std::lock_guard<std::mutex> g(c->ex->mu);
    HIPCHECK(hipSetDevice(c->ex->device));
    WireHost h;
    pull_state(c, h);
    std::string out;
    wire_serialize(WireView{c->m, c->S, (long)c->tick, (long)c->tick0, &c->pos0}, h, out);
    return out;
}

} // namespace kunrt
