// Device runtime of the factor pipeline: executors, module libraries, batch
// graph execution and stream contexts, exported through the C ABI of
// include/kun_capi.h.
//
// Replaces the reference's CPU runtime (cpp/Kun/Runtime.cpp, Executor.cpp,
// Module.cpp, CApi.cpp): instead of a thread pool claiming 8-stock tasks, a
// stage is a short sequence of kernel launches over all stocks (TS stages:
// 64 stocks per workgroup, rows split into chunks where the ops allow;
// cross-sectional stages: one workgroup per row), enqueued in a topological
// order of the stage DAG on the executor's HIP stream.
#include <hip/hip_runtime.h>

#include <dlfcn.h>
#include <stdint.h>
#include <string.h>

#include <algorithm>
#include <map>
#include <memory>
#include <mutex>
#include <stdexcept>
#include <string>
#include <unordered_map>
#include <vector>

#include "../../include/kun_capi.h"
#include "kun_cs.hpp"
#include "kun_device.hpp"
#include "kun_module_abi.h"
#include "kun_rt.hpp"

namespace kunrt {

thread_local std::string g_last_error;

static long env_long(const char *name, long dflt) {
    const char *v = getenv(name);
    return (v && *v) ? atol(v) : dflt;
}

void set_error(const std::string &msg) { g_last_error = msg; }
const char *last_error() { return g_last_error.c_str(); }

// ---------------------------------------------------------------------------
// device buffer pool: exact-size caching allocator (hipMalloc only on miss)
void *Pool::get(size_t bytes) {
    if (bytes == 0) bytes = 256;
    std::lock_guard<std::mutex> g(mu);
    auto it = free_.find(bytes);
    if (it != free_.end()) {
        void *p = it->second;
        free_.erase(it);
        return p;
    }
    void *p = nullptr;
    hipError_t e = hipMalloc(&p, bytes);
    if (e != hipSuccess) {
        // release cached blocks and retry once
        for (auto &kv : free_) hipFree(kv.second);
        free_.clear();
        e = hipMalloc(&p, bytes);
        if (e != hipSuccess) throw Error(std::string("hipMalloc failed: ") + hipGetErrorString(e));
    }
    return p;
}

void Pool::put(void *p, size_t bytes) {
    if (!p) return;
    if (bytes == 0) bytes = 256;
    std::lock_guard<std::mutex> g(mu);
    free_.emplace(bytes, p);
}

Pool::~Pool() {
    for (auto &kv : free_) hipFree(kv.second);
}

Executor::Executor() : chunking(ChunkPolicy::from_env()) {
    HIPCHECK(hipGetDevice(&device));
    HIPCHECK(hipStreamCreateWithFlags(&stream, hipStreamNonBlocking));
    // lanes for independent stages (default priority, like every stream here)
    long nl = env_long("KUN_STAGE_LANES", 4);
    nl = nl < 0 ? 0 : (nl > 8 ? 8 : nl);
    HIPCHECK(hipEventCreateWithFlags(&fork, hipEventDisableTiming));
    for (long l = 0; l < nl; l++) {
        hipStream_t s;
        hipEvent_t e;
        HIPCHECK(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
        HIPCHECK(hipEventCreateWithFlags(&e, hipEventDisableTiming));
        lanes.push_back(s);
        lane_done.push_back(e);
    }
}

void Executor::ensure_stage_events(int n) {
    while ((int)stage_done.size() < n) {
        hipEvent_t e;
        HIPCHECK(hipEventCreateWithFlags(&e, hipEventDisableTiming));
        stage_done.push_back(e);
    }
}

Executor::~Executor() {
    hipStreamSynchronize(stream);
    for (auto l : lanes) hipStreamSynchronize(l);
    for (auto l : lanes) hipStreamDestroy(l);
    for (auto e : lane_done) hipEventDestroy(e);
    for (auto e : stage_done) hipEventDestroy(e);
    if (fork) hipEventDestroy(fork);
    for (auto &p : pending) {
        hipEventDestroy(p.start);
        hipEventDestroy(p.stop);
    }
    for (auto e : event_pool) hipEventDestroy(e);
    free_temps();
    hipStreamDestroy(stream);
}

hipEvent_t Executor::take_event() {
    if (!event_pool.empty()) {
        hipEvent_t e = event_pool.back();
        event_pool.pop_back();
        return e;
    }
    hipEvent_t e;
    HIPCHECK(hipEventCreate(&e));
    return e;
}

void Executor::collect_profile() {
    for (auto &p : pending) {
        HIPCHECK(hipEventSynchronize(p.stop));
        float ms = 0;
        HIPCHECK(hipEventElapsedTime(&ms, p.start, p.stop));
        auto &acc = stage_ms[std::make_tuple(p.module, p.stage, p.kern)];
        acc.first += ms;
        acc.second += 1;
        event_pool.push_back(p.start);
        event_pool.push_back(p.stop);
    }
    pending.clear();
}

// ---------------------------------------------------------------------------
std::string module_symbol(const std::string &name) {
    std::string s = "kun_module_";
    for (char c : name) s += (isalnum((unsigned char)c) ? c : '_');
    return s;
}

Library::~Library() {
    if (handle) dlclose(handle);
}

std::shared_ptr<Library> Library::load(const char *path) {
    void *h = dlopen(path, RTLD_NOW | RTLD_LOCAL);
    if (!h) {
        fprintf(stderr, "dlopen failed: %s %s\n", path, dlerror());
        return nullptr;
    }
    auto lib = std::make_shared<Library>();
    lib->handle = h;
    lib->path = path;
    return lib;
}

const KunModuleDesc *Library::get_module(const char *name) {
    auto *m = (const KunModuleDesc *)dlsym(handle, module_symbol(name).c_str());
    if (m && m->abi != KUN_MODULE_ABI_VERSION) return nullptr;
    return m;
}

// topological order of the stage DAG (cached per descriptor)
static std::mutex g_order_mu;
static std::unordered_map<const KunModuleDesc *, std::vector<int>> g_orders;

const std::vector<int> &stage_order(const KunModuleDesc *m) {
    std::lock_guard<std::mutex> g(g_order_mu);
    auto it = g_orders.find(m);
    if (it != g_orders.end()) return it->second;
    std::vector<int> indeg(m->nstage), order;
    for (int i = 0; i < m->nstage; i++) indeg[i] = m->stages[i].ndeps;
    for (int i = 0; i < m->nstage; i++)
        if (!indeg[i]) order.push_back(i);
    for (size_t h = 0; h < order.size(); h++) {
        int s = order[h];
        for (int j = 0; j < m->nstage; j++)
            for (int d = 0; d < m->stages[j].ndeps; d++)
                if (m->stages[j].deps[d] == s && --indeg[j] == 0) order.push_back(j);
    }
    if ((int)order.size() != m->nstage) throw Error("stage graph of the module has a cycle");
    return g_orders.emplace(m, std::move(order)).first->second;
}

static bool is_single_value(const KunBufferDesc &b) { return strncmp(b.name, "__init", 6) == 0; }

size_t elem_size(const KunModuleDesc *m) { return m->dtype == KUN_DTYPE_DOUBLE ? 8 : 4; }

// elements a buffer holds for a batch run
size_t buffer_elems(const KunModuleDesc *m, int b, size_t S, size_t T, size_t L) {
    const KunBufferDesc &d = m->bufs[b];
    size_t blk = (size_t)m->blocking;
    size_t Sr = (S + blk - 1) / blk * blk;
    if (d.kind == KUN_BUF_INPUT) {
        if (is_single_value(d)) return S;
        return (m->in_layout == KUN_LAYOUT_STS ? Sr : S) * T;
    }
    if (d.kind == KUN_BUF_OUTPUT) return (m->out_layout == KUN_LAYOUT_STS ? Sr : S) * L;
    return S * L;
}

static kun_cs::View cs_view(const KunModuleDesc *m, int b, void *ptr, size_t T, size_t start, size_t L) {
    const KunBufferDesc &d = m->bufs[b];
    kun_cs::View v;
    v.ptr = ptr;
    if (d.kind == KUN_BUF_INPUT) {
        v.layout = m->in_layout == KUN_LAYOUT_STS ? 0 : 1;
        v.rows = (long)T;
        v.row0 = (long)start;
    } else {
        v.layout = (d.kind == KUN_BUF_OUTPUT && m->out_layout == KUN_LAYOUT_STS) ? 0 : 1;
        v.rows = (long)L;
        v.row0 = 0;
    }
    return v;
}

// ---------------------------------------------------------------------------
// Launch geometry and HBM scratch of a batch run.

ChunkPolicy ChunkPolicy::from_env() {
    ChunkPolicy p;
    p.target_workgroups = env_long("KUN_TARGET_WORKGROUPS", p.target_workgroups);
    p.halo_factor = env_long("KUN_HALO_FACTOR", p.halo_factor);
    p.min_rows = env_long("KUN_MIN_CHUNK_ROWS", p.min_rows);
    return p;
}

// Row chunks of a PAR kernel: enough workgroups to fill the chip, each chunk
// at least `halo_factor` times the halo it recomputes and `min_rows` long.
static long pick_chunks(const KunKernelDesc &kd, size_t S, size_t L, const ChunkPolicy &pol) {
    if (kd.kind != KUN_KERNEL_PAR || L == 0) return 1;
    long nsb = ((long)S + 63) / 64;
    long rows_min = std::max(std::max(1L, pol.min_rows), pol.halo_factor * (long)kd.halo);
    long max_ch = std::max(1L, (long)L / rows_min);
    long want = std::max(1L, (pol.target_workgroups + nsb * kd.ncomp - 1) / (nsb * kd.ncomp));
    long nch = std::min(max_ch, want);
    long chunk = ((long)L + nch - 1) / nch;
    return ((long)L + chunk - 1) / chunk;
}

static size_t align256(size_t x) { return (x + 255) & ~(size_t)255; }

BatchLayout batch_layout(const KunModuleDesc *m, size_t S, size_t L, const ChunkPolicy &pol) {
    BatchLayout lay;
    const size_t esz = elem_size(m);
    lay.stage_mat.assign(m->nstage, 0);
    lay.kern.resize(m->nstage);
    size_t off = 0;
    for (int i = 0; i < m->nstage; i++) {
        const KunStageDesc &sd = m->stages[i];
        if (sd.kind != KUN_STAGE_TS) continue;
        if (sd.nmat) {
            lay.stage_mat[i] = off;
            off += align256((size_t)sd.nmat * L * S * esz);
        }
        for (int j = 0; j < sd.nkern; j++) {
            BatchLayout::Kern k;
            k.chunks = pick_chunks(sd.kerns[j], S, L, pol);
            k.chunk_rows = L ? ((long)L + k.chunks - 1) / k.chunks : 1;
            k.scratch = off;
            off += align256((size_t)sd.kerns[j].spill_bytes * S * k.chunks);
            lay.kern[i].push_back(k);
        }
    }
    lay.bytes = off;
    return lay;
}

// Enqueue one stage (all its kernels) on stream `st`.
static void enqueue_stage(const KunModuleDesc *m, int si, const std::vector<void *> &dev, const BatchLayout &lay,
                          char *arena, size_t S, size_t T, size_t start, size_t L, hipStream_t st, Executor *prof) {
    const KunStageDesc &sd = m->stages[si];
    int nk = sd.kind == KUN_STAGE_TS ? sd.nkern : 1;
    for (int kj = 0; kj < nk; kj++) {
        Executor::Pending ev{m, si, kj, nullptr, nullptr};
        if (prof) {
            ev.start = prof->take_event();
            ev.stop = prof->take_event();
            HIPCHECK(hipEventRecord(ev.start, st));
        }
        if (sd.kind == KUN_STAGE_TS) {
            const KunKernelDesc &kd = sd.kerns[kj];
            const BatchLayout::Kern &kl = lay.kern[si][kj];
            kun::StageArgs a;
            memset(&a, 0, sizeof a);
            int n = 0;
            for (int j = 0; j < sd.nin; j++) a.bufs[n++] = dev[sd.in[j]];
            for (int j = 0; j < sd.nout; j++) a.bufs[n++] = dev[sd.out[j]];
            a.S = (long)S;
            a.T = (long)T;
            a.start = (long)start;
            a.L = (long)L;
            a.chunk = kl.chunk_rows;
            a.blk = m->blocking;
            a.scratch = kd.spill_bytes ? arena + kl.scratch : nullptr;
            a.mat = sd.nmat ? arena + lay.stage_mat[si] : nullptr;
            void *args[] = {&a};
            dim3 grid((unsigned)((S + sd.block - 1) / sd.block), (unsigned)kl.chunks, (unsigned)kd.ncomp);
            HIPCHECK(hipLaunchKernel(kd.kernel, grid, dim3(sd.block), args, 0, st));
        } else {
            int ib = sd.in[0], ob = sd.out[0];
            kun_cs::View vi = cs_view(m, ib, dev[ib], T, start, L);
            kun_cs::View vo = cs_view(m, ob, dev[ob], T, start, L);
            if (sd.kind == KUN_STAGE_RANK && (long)S > kun_cs::max_rank_stocks(m->dtype))
                throw Error("Rank over more than " + std::to_string(kun_cs::max_rank_stocks(m->dtype)) +
                            " stocks is not supported");
            hipError_t e = sd.kind == KUN_STAGE_RANK
                               ? kun_cs::launch_rank(m->dtype, vi, vo, (long)S, (long)L, m->blocking, st)
                               : kun_cs::launch_scale(m->dtype, vi, vo, (long)S, (long)L, m->blocking, st);
            if (e != hipSuccess) throw Error(std::string("cross-sectional kernel: ") + hipGetErrorString(e));
        }
        if (prof) {
            HIPCHECK(hipEventRecord(ev.stop, st));
            prof->pending.push_back(ev);
        }
    }
}

// Enqueue every stage of a batch module; dev[b] are device pointers, arena
// the scratch described by `lay`. Stages run on the executor's lanes (HIP
// streams) in a topological order, round robin; a stage waits for the
// stages it reads from through events, and `st` waits for every lane at the
// end (and every lane for `st` at the start).
void enqueue_batch(Executor &ex, const KunModuleDesc *m, const std::vector<void *> &dev, const BatchLayout &lay,
                   char *arena, size_t S, size_t T, size_t start, size_t L, hipStream_t st) {
    if (L == 0) return;
    Executor *prof = ex.profiling ? &ex : nullptr;
    const std::vector<int> &order = stage_order(m);
    const int nl = (int)ex.lanes.size();
    if (nl == 0 || m->nstage == 1) {
        for (int si : order) enqueue_stage(m, si, dev, lay, arena, S, T, start, L, st, prof);
        return;
    }
    ex.ensure_stage_events(m->nstage);
    HIPCHECK(hipEventRecord(ex.fork, st));
    for (int l = 0; l < nl; l++) HIPCHECK(hipStreamWaitEvent(ex.lanes[l], ex.fork, 0));
    std::vector<int> lane_of(m->nstage, 0);
    int next = 0;
    for (int si : order) {
        const KunStageDesc &sd = m->stages[si];
        const int l = next;
        next = (next + 1) % nl;
        lane_of[si] = l;
        hipStream_t ls = ex.lanes[l];
        for (int d = 0; d < sd.ndeps; d++) {
            int dep = sd.deps[d];
            if (lane_of[dep] != l) HIPCHECK(hipStreamWaitEvent(ls, ex.stage_done[dep], 0));
        }
        enqueue_stage(m, si, dev, lay, arena, S, T, start, L, ls, prof);
        HIPCHECK(hipEventRecord(ex.stage_done[si], ls));
    }
    for (int l = 0; l < nl; l++) {
        HIPCHECK(hipEventRecord(ex.lane_done[l], ex.lanes[l]));
        HIPCHECK(hipStreamWaitEvent(st, ex.lane_done[l], 0));
    }
}

void check_batch_module(const KunModuleDesc *m, size_t S, size_t T, size_t start, size_t L) {
    if (!m) throw Error("null module");
    if (m->abi != KUN_MODULE_ABI_VERSION)
        throw Error("The required version in the module does not match the runtime version");
    if (m->out_layout == KUN_LAYOUT_STREAM) throw Error("Cannot run stream mode module via runGraph()");
    if (start + L > T) throw Error("Bad parameter: cur_time + length exceeds total_time");
    if (m->aligned && S % (size_t)m->blocking) throw Error("num_stocks must be a multiple of blocking_len");
}

// Host-buffer run: H2D inputs, enqueue, D2H outputs, synchronise.
void run_graph_host(Executor &ex, const KunModuleDesc *m, const NameMap &bufs, size_t S, size_t T,
                    size_t start, size_t L) {
    check_batch_module(m, S, T, start, L);
    std::lock_guard<std::mutex> g(ex.mu);
    HIPCHECK(hipSetDevice(ex.device));
    const size_t esz = elem_size(m);
    std::vector<void *> dev(m->nbuf, nullptr);
    std::vector<size_t> bytes(m->nbuf, 0);
    struct Release {
        Executor &ex;
        std::vector<void *> &dev;
        std::vector<size_t> &bytes;
        ~Release() {
            hipStreamSynchronize(ex.stream);
            for (size_t i = 0; i < dev.size(); i++) ex.pool.put(dev[i], bytes[i]);
        }
    } release{ex, dev, bytes};
    for (int b = 0; b < m->nbuf; b++) {
        const KunBufferDesc &d = m->bufs[b];
        bytes[b] = buffer_elems(m, b, S, T, L) * esz;
        if (d.kind != KUN_BUF_TEMP && !bufs.count(d.name))
            throw Error(std::string("Buffer name not found: ") + d.name);
        dev[b] = ex.pool.get(bytes[b]);
        if (d.kind == KUN_BUF_INPUT)
            HIPCHECK(hipMemcpyAsync(dev[b], bufs.at(d.name), bytes[b], hipMemcpyHostToDevice, ex.stream));
    }
    BatchLayout lay = batch_layout(m, S, L, ex.chunking);
    struct ReleaseArena {
        Executor &ex;
        char *p = nullptr;
        size_t bytes;
        ~ReleaseArena() {
            hipStreamSynchronize(ex.stream);
            ex.pool.put(p, bytes);
        }
    } arena{ex, nullptr, lay.bytes};
    if (lay.bytes) arena.p = (char *)ex.pool.get(lay.bytes);
    enqueue_batch(ex, m, dev, lay, arena.p, S, T, start, L, ex.stream);
    for (int b = 0; b < m->nbuf; b++) {
        const KunBufferDesc &d = m->bufs[b];
        if (d.kind == KUN_BUF_OUTPUT)
            HIPCHECK(hipMemcpyAsync(bufs.at(d.name), dev[b], bytes[b], hipMemcpyDeviceToHost, ex.stream));
    }
    HIPCHECK(hipStreamSynchronize(ex.stream));
}

// Device-buffer run: temporaries from the pool (kept for reuse per shape).
void run_graph_device(Executor &ex, const KunModuleDesc *m, const NameMap &bufs, size_t S, size_t T,
                      size_t start, size_t L, hipStream_t st) {
    check_batch_module(m, S, T, start, L);
    std::lock_guard<std::mutex> g(ex.mu);
    HIPCHECK(hipSetDevice(ex.device));
    if (!st) st = ex.stream;
    const size_t esz = elem_size(m);
    auto key = std::make_tuple((const void *)m, S, L);
    auto it = ex.temps.find(key);
    if (it == ex.temps.end()) {
        // temporaries first, then the scratch arena
        DeviceTemps t;
        t.lay = batch_layout(m, S, L, ex.chunking);
        t.bufs.assign(m->nbuf, nullptr);
        for (int b = 0; b < m->nbuf; b++)
            if (m->bufs[b].kind == KUN_BUF_TEMP) {
                size_t nb = buffer_elems(m, b, S, T, L) * esz;
                HIPCHECK(hipMalloc(&t.bufs[b], nb ? nb : 256));
            }
        if (t.lay.bytes) HIPCHECK(hipMalloc((void **)&t.arena, t.lay.bytes));
        it = ex.temps.emplace(key, std::move(t)).first;
    }
    const DeviceTemps &temps = it->second;
    std::vector<void *> dev(m->nbuf, nullptr);
    for (int b = 0; b < m->nbuf; b++) {
        const KunBufferDesc &d = m->bufs[b];
        if (d.kind == KUN_BUF_TEMP) {
            dev[b] = temps.bufs[b];
        } else {
            auto f = bufs.find(d.name);
            if (f == bufs.end()) throw Error(std::string("Buffer name not found: ") + d.name);
            dev[b] = f->second;
        }
    }
    enqueue_batch(ex, m, dev, temps.lay, temps.arena, S, T, start, L, st);
}

void Executor::free_temps() {
    for (auto &kv : temps) {
        for (void *p : kv.second.bufs)
            if (p) hipFree(p);
        if (kv.second.arena) hipFree(kv.second.arena);
    }
    temps.clear();
}

} // namespace kunrt
