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
// Serialised state (own format, not byte-compatible with the reference):
//   "KUNSTRM1" | u64 num_stocks | u64 nbuf | u64 nstage | u64 tick
//   | per buffer: row bytes | per stage: u64 bytes + state bytes
// Short input while reading the rows raises "Failed to read initial stream
// buffer", short input in the stage section "Failed to deserialize state
// buffer" (the reference's messages, Runtime.cpp:404-406, 454-458).
#include <hip/hip_runtime.h>
#include <string.h>

#include <fstream>
#include <iterator>
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
    if (state_src) {
        std::string blob;
        if (from_file) {
            std::ifstream f(state_src, std::ios::binary);
            blob.assign(std::istreambuf_iterator<char>(f), std::istreambuf_iterator<char>());
        } else {
            blob.assign(state_src, state_len);
        }
        size_t pos = 0;
        auto take = [&](void *dst, size_t n) {
            if (pos + n > blob.size()) return false;
            memcpy(dst, blob.data() + pos, n);
            pos += n;
            return true;
        };
        char magic[8];
        uint64_t hdr[4];
        if (!take(magic, 8) || memcmp(magic, "KUNSTRM1", 8) || !take(hdr, sizeof hdr) || hdr[0] != S ||
            hdr[1] != (uint64_t)m->nbuf || hdr[2] != (uint64_t)m->nstage)
            throw Error("Failed to read initial stream buffer");
        std::vector<char> tmp(c->esz * S);
        for (int b = 0; b < m->nbuf; b++) {
            if (!take(tmp.data(), tmp.size())) throw Error("Failed to read initial stream buffer");
            HIPCHECK(hipMemcpy(c->rows[b], tmp.data(), tmp.size(), hipMemcpyHostToDevice));
        }
        for (int si = 0; si < m->nstage; si++) {
            uint64_t n = 0;
            if (!take(&n, 8) || n != c->state_len[si]) throw Error("Failed to deserialize state buffer");
            std::vector<char> st(n);
            if (n && !take(st.data(), n)) throw Error("Failed to deserialize state buffer");
            if (n) HIPCHECK(hipMemcpy(c->state[si], st.data(), n, hipMemcpyHostToDevice));
        }
        c->tick = hdr[3];
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
    const KunModuleDesc *m = c->m;
    std::lock_guard<std::mutex> g(c->ex->mu);
    HIPCHECK(hipStreamSynchronize(c->ex->stream));
    std::string out("KUNSTRM1", 8);
    uint64_t hdr[4] = {c->S, (uint64_t)m->nbuf, (uint64_t)m->nstage, c->tick};
    out.append((const char *)hdr, sizeof hdr);
    std::vector<char> tmp(c->esz * c->S);
    for (int b = 0; b < m->nbuf; b++) {
        HIPCHECK(hipMemcpy(tmp.data(), c->rows[b], tmp.size(), hipMemcpyDeviceToHost));
        out.append(tmp.data(), tmp.size());
    }
    for (int si = 0; si < m->nstage; si++) {
        uint64_t n = c->state_len[si];
        out.append((const char *)&n, 8);
        if (n) {
            std::vector<char> st(n);
            HIPCHECK(hipMemcpy(st.data(), c->state[si], n, hipMemcpyDeviceToHost));
            out.append(st.data(), n);
        }
    }
    return out;
}

} // namespace kunrt
