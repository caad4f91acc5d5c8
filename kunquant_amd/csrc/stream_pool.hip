// Session pools: many stream sessions of one STREAM module resident in HBM
// side by side, ticked together.
//
// A pool owns one stock space of `capacity` lanes. Every module buffer is one
// row of `capacity` values and every time-series stage owns its carried
// state (history rings, op state) as [slot][capacity] arrays, exactly the
// layout of a single stream context with num_stocks = capacity. A session is
// a range [s0, s0 + S) of that space (64-lane aligned), taken first-fit at
// admission and returned at release, so the pool admits sessions by their
// state footprint.
//
// A cycle ticks any set of idle sessions once, each with its own tick
// counter: one host-to-device copy of the cycle's tables and packed input
// rows, one scatter kernel, then per phase of the module's launch plan
// (KunPhaseDesc) one time-series kernel over the blocks of every session in
// the cycle, or one Rank and one Scale launch with a workgroup per (stage,
// session). Results equal those of one stream context per session: the same
// per-stock tick bodies and row kernels run on the same values.
//
// Cycles run on the pool's own streams; up to `max_inflight` cycles with
// disjoint sessions are on the device at once (a session belongs to at most
// one cycle in flight), so ticks that arrive while a cycle runs start
// without waiting for it.
//
// Readback: with a set of readback buffers (pool_set_readback), a cycle ends
// by gathering those rows of its sessions into one packed device block
// ([buffer][lanes of the cycle]) and copying it to the cycle slot's pinned
// host block, so that a finished cycle's results are already in host memory
// (pool_result points into that block until the slot's next cycle).
// Profiling (pool_profile) records events between the input staging, the
// phases and the readback of each cycle (pool_phase_times).
#include <hip/hip_runtime.h>
#include <string.h>

#include <condition_variable>
#include <functional>
#include <map>
#include <string>
#include <thread>
#include <vector>

#include "kun_cs.hpp"
#include "kun_device.hpp"
#include "kun_module_abi.h"
#include "kun_rt.hpp"

namespace kunrt {

namespace {

constexpr int POOL_MAX_INPUTS = 16;

struct ScatterArgs {
    void *dst[POOL_MAX_INPUTS]; // pool rows of the module's inputs
    const void *packed;         // [input][total] rows of the cycle's sessions
    long total;
    const kun::PoolBlock *blocks;
    int nin;
};

template <typename T> __global__ __launch_bounds__(64) void pool_scatter(ScatterArgs a) {
    const kun::PoolBlock e = a.blocks[blockIdx.x];
    const long s = (long)e.s0 + threadIdx.x;
    if (s >= e.s1) return;
    const long k = e.p0 + threadIdx.x;
    for (int q = 0; q < a.nin; q++) ((T *)a.dst[q])[s] = ((const T *)a.packed)[(size_t)q * a.total + k];
}

struct GatherArgs {
    const void *const *src; // pool rows of the readback buffers
    void *packed;           // [buffer][total] results of the cycle's sessions
    long total;
    const kun::PoolBlock *blocks;
    int nq;
};

template <typename T> __global__ __launch_bounds__(64) void pool_gather(GatherArgs a) {
    const kun::PoolBlock e = a.blocks[blockIdx.x];
    const long s = (long)e.s0 + threadIdx.x;
    if (s >= e.s1) return;
    const long k = e.p0 + threadIdx.x;
    for (int q = 0; q < a.nq; q++) ((T *)a.packed)[(size_t)q * a.total + k] = ((const T *)a.src[q])[s];
}

// NaN into [s0, s0 + n) of `nrow` rows of `stride` values
template <typename T> __global__ void pool_fill_nan(T *rows, long stride, long s0, long n) {
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) rows[(size_t)blockIdx.y * stride + s0 + i] = kun::qnan<T>();
}

size_t align256(size_t x) { return (x + 255) & ~(size_t)255; }

} // namespace

// Host threads that copy a cycle's input rows into pinned staging: a large
// cycle moves megabytes of rows, more than one core copies in the time the
// device takes for a cycle.
class CopyWorkers {
  public:
    explicit CopyWorkers(int n) {
        for (int i = 0; i < n; i++) threads_.emplace_back([this] { loop(); });
    }
    ~CopyWorkers() {
        {
            std::lock_guard<std::mutex> g(mu_);
            stop_ = true;
        }
        cv_.notify_all();
        for (auto &t : threads_) t.join();
    }
    // runs fn(i) for i in [0, n), the calling thread included
    void run(size_t n, const std::function<void(size_t)> &fn) {
        {
            std::lock_guard<std::mutex> g(mu_);
            fn_ = &fn;
            n_ = n;
            next_ = 0;
            active_ = (int)threads_.size();
            gen_++;
        }
        cv_.notify_all();
        work();
        std::unique_lock<std::mutex> g(mu_);
        done_cv_.wait(g, [this] { return active_ == 0; });
        fn_ = nullptr;
    }
    int size() const { return (int)threads_.size(); }

  private:
    void work() {
        for (;;) {
            size_t i;
            {
                std::lock_guard<std::mutex> g(mu_);
                if (next_ >= n_) return;
                i = next_++;
            }
            (*fn_)(i);
        }
    }
    void loop() {
        unsigned long seen = 0;
        for (;;) {
            {
                std::unique_lock<std::mutex> g(mu_);
                cv_.wait(g, [&] { return stop_ || gen_ != seen; });
                if (stop_) return;
                seen = gen_;
            }
            work();
            std::lock_guard<std::mutex> g(mu_);
            if (--active_ == 0) done_cv_.notify_all();
        }
    }
    std::vector<std::thread> threads_;
    std::mutex mu_;
    std::condition_variable cv_, done_cv_;
    const std::function<void(size_t)> *fn_ = nullptr;
    size_t n_ = 0, next_ = 0;
    int active_ = 0;
    unsigned long gen_ = 0;
    bool stop_ = false;
};

struct PoolSession {
    bool live = false;
    size_t s0 = 0, S = 0;
    long tick = 0;
    int slot = -1; // in-flight cycle slot, or -1
    // reference wire positions recorded when the session resumed from a
    // serialised state (see WireView)
    long tick0 = 0;
    std::vector<uint64_t> pos0;
};

struct PoolSlot {
    hipStream_t st = nullptr;
    hipEvent_t done = nullptr;
    hipEvent_t computed = nullptr; // phases (and the readback gather) finished
    char *host = nullptr; // pinned staging
    char *dev = nullptr;
    bool busy = false;
    unsigned long long cycle = 0;
    std::vector<int> sessions;
    // readback: the cycle's results, [buffer][res_total] (pinned, device)
    char *res_host = nullptr, *res_dev = nullptr;
    size_t res_total = 0;
    std::map<int, size_t> res_off; // session -> first lane in the block
    // profiling: events at the cycle's boundaries (timing enabled)
    std::vector<hipEvent_t> ph_ev;
    bool profiled = false;
};

struct StreamPool {
    std::shared_ptr<Executor> ex;
    const KunModuleDesc *m = nullptr;
    size_t cap = 0, esz = 4;
    int max_sessions = 0;
    char *rows = nullptr;                 // [nbuf][cap]
    std::vector<void *> state;            // per stage [state_bytes][cap]
    kun::StreamArgs *d_args = nullptr;    // per stage
    std::vector<int> inputs;              // input buffer indices, module order
    std::vector<int> ncs_rank, ncs_scale; // per phase: Rank / Scale stages
    size_t stage_bytes = 0;               // per-stock bytes of rows + state
    std::map<size_t, size_t> free_;       // start -> length, 64-lane units
    std::vector<PoolSession> sess;
    std::vector<PoolSlot> slots;
    size_t staging = 0;
    unsigned long long next_cycle = 1;
    char *scratch = nullptr;              // admission block table
    std::unique_ptr<CopyWorkers> copiers;
    std::vector<int> rb_bufs;             // readback buffer handles
    const void **d_rb_src = nullptr;      // their pool rows (device array)
    bool profile = false;
    ~StreamPool();
    void *row(int b) const { return rows + (size_t)b * cap * esz; }
};

StreamPool::~StreamPool() {
    if (!ex) return;
    hipSetDevice(ex->device);
    for (auto &sl : slots) {
        if (sl.st) hipStreamSynchronize(sl.st);
        if (sl.st) hipStreamDestroy(sl.st);
        if (sl.done) hipEventDestroy(sl.done);
        if (sl.computed) hipEventDestroy(sl.computed);
        if (sl.host) hipHostFree(sl.host);
        if (sl.dev) hipFree(sl.dev);
        if (sl.res_host) hipHostFree(sl.res_host);
        if (sl.res_dev) hipFree(sl.res_dev);
        for (auto e : sl.ph_ev) hipEventDestroy(e);
    }
    if (d_rb_src) hipFree(d_rb_src);
    hipStreamSynchronize(ex->stream);
    if (rows) hipFree(rows);
    for (void *p : state)
        if (p) hipFree(p);
    if (d_args) hipFree(d_args);
    if (scratch) hipFree(scratch);
}

size_t pool_bytes_per_stock(const KunModuleDesc *m) {
    size_t b = (size_t)m->nbuf * elem_size(m);
    for (int si = 0; si < m->nstage; si++)
        if (m->stages[si].kind == KUN_STAGE_TS) b += (size_t)m->stages[si].state_bytes;
    return b;
}

static void retire(StreamPool *p, PoolSlot &sl) {
    for (int sid : sl.sessions) p->sess[sid].slot = -1;
    sl.sessions.clear();
    sl.busy = false;
}

static void wait_slot(StreamPool *p, PoolSlot &sl) {
    if (!sl.busy) return;
    HIPCHECK(hipEventSynchronize(sl.done));
    retire(p, sl);
}

StreamPool *pool_create(std::shared_ptr<Executor> ex, const KunModuleDesc *m, size_t capacity, int max_sessions,
                        int max_inflight) {
    if (!m || m->abi != KUN_MODULE_ABI_VERSION)
        throw Error("The required version in the module does not match the runtime version");
    if (m->out_layout != KUN_LAYOUT_STREAM || !m->phases) throw Error("Session pools need a STREAM module");
    if (capacity == 0 || max_sessions < 1 || max_inflight < 1) throw Error("bad session pool parameters");
    std::unique_ptr<StreamPool> p(new StreamPool());
    p->ex = ex;
    p->m = m;
    p->cap = (capacity + 63) / 64 * 64;
    if (p->cap > 0x7fffffffUL) throw Error("session pool capacity too large");
    p->esz = elem_size(m);
    p->max_sessions = max_sessions;
    for (int b = 0; b < m->nbuf; b++)
        if (m->bufs[b].kind == KUN_BUF_INPUT) p->inputs.push_back(b);
    if ((int)p->inputs.size() > POOL_MAX_INPUTS) throw Error("session pools take at most 16 inputs");
    for (int si = 0; si < m->nstage; si++)
        if (m->stages[si].kind == KUN_STAGE_GENERIC)
            throw Error("session pools do not run generic cross-sectional stages; use a stream context");
    size_t max_segs = 0;
    for (int ph = 0; ph < m->nphase; ph++) {
        const KunPhaseDesc &pd = m->phases[ph];
        int nr = 0, ns = 0;
        if (pd.kind == KUN_PHASE_CS)
            for (int j = 0; j < pd.nstages; j++) (m->stages[pd.stages[j]].kind == KUN_STAGE_RANK ? nr : ns)++;
        p->ncs_rank.push_back(nr);
        p->ncs_scale.push_back(ns);
        max_segs += (size_t)(nr + ns) * max_sessions;
    }
    p->stage_bytes = pool_bytes_per_stock(m);
    std::lock_guard<std::mutex> g(ex->mu);
    HIPCHECK(hipSetDevice(ex->device));
    HIPCHECK(hipMalloc((void **)&p->rows, (size_t)m->nbuf * p->cap * p->esz));
    p->state.assign(m->nstage, nullptr);
    std::vector<kun::StreamArgs> args(m->nstage);
    for (int si = 0; si < m->nstage; si++) {
        const KunStageDesc &sd = m->stages[si];
        kun::StreamArgs &a = args[si];
        memset(&a, 0, sizeof a);
        if (sd.kind != KUN_STAGE_TS) continue;
        int n = 0;
        for (int j = 0; j < sd.nin; j++) a.bufs[n++] = p->row(sd.in[j]);
        for (int j = 0; j < sd.nout; j++) a.bufs[n++] = p->row(sd.out[j]);
        if (sd.state_bytes) HIPCHECK(hipMalloc(&p->state[si], (size_t)sd.state_bytes * p->cap));
        a.st = p->state[si];
        a.S = (long)p->cap;
    }
    HIPCHECK(hipMalloc((void **)&p->d_args, sizeof(kun::StreamArgs) * m->nstage));
    HIPCHECK(hipMemcpy(p->d_args, args.data(), sizeof(kun::StreamArgs) * m->nstage, hipMemcpyHostToDevice));
    // staging: block table, segment tables, packed input rows
    const size_t max_blocks = p->cap / 64;
    p->staging = align256(max_blocks * sizeof(kun::PoolBlock)) + align256(max_segs * sizeof(kun_cs::Seg) + 1) +
                 align256(p->inputs.size() * p->cap * p->esz + 1);
    p->slots.resize(max_inflight);
    for (auto &sl : p->slots) {
        HIPCHECK(hipStreamCreateWithFlags(&sl.st, hipStreamNonBlocking));
        HIPCHECK(hipEventCreateWithFlags(&sl.done, hipEventDisableTiming));
        HIPCHECK(hipEventCreateWithFlags(&sl.computed, hipEventDisableTiming));
        HIPCHECK(hipHostMalloc((void **)&sl.host, p->staging, hipHostMallocDefault));
        HIPCHECK(hipMalloc((void **)&sl.dev, p->staging));
    }
    HIPCHECK(hipMalloc((void **)&p->scratch, align256(max_blocks * sizeof(kun::PoolBlock))));
    p->free_[0] = p->cap / 64;
    p->copiers.reset(new CopyWorkers(7));
    return p.release();
}

void pool_destroy(StreamPool *p) { delete p; }

size_t pool_free_stocks(StreamPool *p) {
    size_t n = 0;
    for (auto &kv : p->free_) n += kv.second;
    return n * 64;
}

// -1: no contiguous room (the caller retries after a release)
int pool_admit(StreamPool *p, size_t S) {
    if (S == 0) throw Error("a session needs at least one stock");
    const KunModuleDesc *m = p->m;
    if (m->aligned && S % (size_t)m->blocking) throw Error("num_stocks must be a multiple of blocking_len");
    for (int si = 0; si < m->nstage; si++)
        if (m->stages[si].kind == KUN_STAGE_RANK && (long)S > kun_cs::max_rank_stocks(m->dtype))
            throw Error("Rank over more than " + std::to_string(kun_cs::max_rank_stocks(m->dtype)) +
                        " stocks is not supported");
    const size_t units = (S + 63) / 64;
    auto it = p->free_.begin();
    while (it != p->free_.end() && it->second < units) ++it;
    if (it == p->free_.end()) return -1;
    int sid = -1;
    for (size_t i = 0; i < p->sess.size(); i++)
        if (!p->sess[i].live) {
            sid = (int)i;
            break;
        }
    if (sid < 0) {
        if ((int)p->sess.size() >= p->max_sessions) return -1;
        sid = (int)p->sess.size();
        p->sess.emplace_back();
    }
    const size_t start = it->first, len = it->second;
    p->free_.erase(it);
    if (len > units) p->free_[start + units] = len - units;
    PoolSession &ss = p->sess[sid];
    ss.live = true;
    ss.s0 = start * 64;
    ss.S = S;
    ss.tick = 0;
    ss.slot = -1;
    ss.tick0 = 0;
    ss.pos0.assign(m->nwire_buf, 0);
    // NaN rows and initial state over the session's lanes (the pool's
    // streams touch other sessions only)
    std::lock_guard<std::mutex> g(p->ex->mu);
    HIPCHECK(hipSetDevice(p->ex->device));
    hipStream_t st = p->ex->stream;
    dim3 grid((unsigned)((S + 255) / 256), (unsigned)m->nbuf);
    if (p->esz == 8)
        hipLaunchKernelGGL(pool_fill_nan<double>, grid, dim3(256), 0, st, (double *)p->rows, (long)p->cap,
                           (long)ss.s0, (long)S);
    else
        hipLaunchKernelGGL(pool_fill_nan<float>, grid, dim3(256), 0, st, (float *)p->rows, (long)p->cap,
                           (long)ss.s0, (long)S);
    HIPCHECK(hipGetLastError());
    std::vector<kun::PoolBlock> blocks;
    for (size_t b = 0; b < units; b++) {
        kun::PoolBlock e;
        e.s0 = (int)(ss.s0 + b * 64);
        e.s1 = (int)std::min(ss.s0 + b * 64 + 64, ss.s0 + S);
        e.p0 = 0;
        e.tick = 0;
        blocks.push_back(e);
    }
    HIPCHECK(hipMemcpyAsync(p->scratch, blocks.data(), blocks.size() * sizeof(kun::PoolBlock), hipMemcpyHostToDevice,
                            st));
    kun::PoolArgs pa{p->d_args, (const kun::PoolBlock *)p->scratch};
    void *args[] = {&pa};
    HIPCHECK(hipLaunchKernel(m->pool_init, dim3((unsigned)units), dim3(64), args, 0, st));
    HIPCHECK(hipStreamSynchronize(st));
    return sid;
}

static PoolSession &live_session(StreamPool *p, int sid) {
    if (sid < 0 || sid >= (int)p->sess.size() || !p->sess[sid].live) throw Error("no such session");
    return p->sess[sid];
}

void pool_release(StreamPool *p, int sid) {
    PoolSession &ss = live_session(p, sid);
    if (ss.slot >= 0) wait_slot(p, p->slots[ss.slot]);
    size_t start = ss.s0 / 64, len = (ss.S + 63) / 64;
    ss.live = false;
    auto nx = p->free_.lower_bound(start);
    if (nx != p->free_.end() && nx->first == start + len) {
        len += nx->second;
        p->free_.erase(nx);
    }
    auto pv = p->free_.lower_bound(start);
    if (pv != p->free_.begin()) {
        --pv;
        if (pv->first + pv->second == start) {
            start = pv->first;
            len += pv->second;
            p->free_.erase(pv);
        }
    }
    p->free_[start] = len;
}

// inputs[j * nin + q]: host row (S values) of input q of sessions[j]
unsigned long long pool_submit(StreamPool *p, int n, const int *sessions, const void *const *inputs) {
    const KunModuleDesc *m = p->m;
    if (n < 1) throw Error("a cycle needs at least one session");
    // the free slot that ran the oldest cycle (so a cycle's readback block
    // outlives the next max_inflight - 1 submissions); the oldest busy one
    // is waited for if none is free
    PoolSlot *sl = nullptr;
    for (auto &x : p->slots) {
        if (x.busy && hipEventQuery(x.done) == hipSuccess) retire(p, x);
        if (!x.busy && (!sl || x.cycle < sl->cycle)) sl = &x;
    }
    if (!sl) {
        sl = &p->slots[0];
        for (auto &x : p->slots)
            if (x.cycle < sl->cycle) sl = &x;
        wait_slot(p, *sl);
    }
    std::vector<char> seen(p->sess.size(), 0);
    size_t total = 0, nblocks = 0;
    long maxS = 0;
    for (int j = 0; j < n; j++) {
        PoolSession &ss = live_session(p, sessions[j]);
        if (ss.slot >= 0) throw Error("session " + std::to_string(sessions[j]) + " is in a cycle in flight");
        if (seen[sessions[j]]) throw Error("session " + std::to_string(sessions[j]) + " twice in one cycle");
        seen[sessions[j]] = 1;
        total += ss.S;
        nblocks += (ss.S + 63) / 64;
        maxS = std::max(maxS, (long)ss.S);
    }
    const size_t nin = p->inputs.size();
    // staging layout
    char *h = sl->host;
    const size_t off_blocks = 0;
    const size_t off_segs = align256(nblocks * sizeof(kun::PoolBlock));
    size_t nseg = 0;
    for (int ph = 0; ph < m->nphase; ph++) nseg += (size_t)(p->ncs_rank[ph] + p->ncs_scale[ph]) * n;
    const size_t off_in = off_segs + align256(nseg * sizeof(kun_cs::Seg) + 1);
    const size_t used = off_in + nin * total * p->esz;
    if (used > p->staging) throw Error("session pool staging overflow");
    kun::PoolBlock *blk = (kun::PoolBlock *)(h + off_blocks);
    std::vector<size_t> offs(n);
    size_t off = 0, bi = 0;
    for (int j = 0; j < n; j++) {
        const PoolSession &ss = p->sess[sessions[j]];
        for (size_t b = 0; b * 64 < ss.S; b++, bi++) {
            blk[bi].s0 = (int)(ss.s0 + b * 64);
            blk[bi].s1 = (int)std::min(ss.s0 + b * 64 + 64, ss.s0 + ss.S);
            blk[bi].p0 = (long)(off + b * 64);
            blk[bi].tick = ss.tick;
        }
        for (size_t q = 0; q < nin; q++)
            if (!inputs[(size_t)j * nin + q]) throw Error("missing input row");
        offs[j] = off;
        off += ss.S;
    }
    auto copy_one = [&](size_t i) {
        const size_t j = i / nin, q = i % nin;
        memcpy(h + off_in + (q * total + offs[j]) * p->esz, inputs[i], p->sess[sessions[j]].S * p->esz);
    };
    if (total * nin * p->esz >= (1u << 20))
        p->copiers->run((size_t)n * nin, copy_one);
    else
        for (size_t i = 0; i < (size_t)n * nin; i++) copy_one(i);
    kun_cs::Seg *seg = (kun_cs::Seg *)(h + off_segs);
    size_t si = 0;
    for (int ph = 0; ph < m->nphase; ph++) {
        const KunPhaseDesc &pd = m->phases[ph];
        if (pd.kind != KUN_PHASE_CS) continue;
        for (int kind : {KUN_STAGE_RANK, KUN_STAGE_SCALE})
            for (int k = 0; k < pd.nstages; k++) {
                const KunStageDesc &sd = m->stages[pd.stages[k]];
                if (sd.kind != kind) continue;
                for (int j = 0; j < n; j++) {
                    const PoolSession &ss = p->sess[sessions[j]];
                    seg[si].in = (const char *)p->row(sd.in[0]) + ss.s0 * p->esz;
                    seg[si].out = (char *)p->row(sd.out[0]) + ss.s0 * p->esz;
                    seg[si].n = (long)ss.S;
                    si++;
                }
            }
    }
    hipStream_t st = sl->st;
    HIPCHECK(hipSetDevice(p->ex->device));
    sl->profiled = p->profile;
    if (p->profile)
        while (sl->ph_ev.size() < (size_t)m->nphase + 3) {
            hipEvent_t e;
            HIPCHECK(hipEventCreate(&e));
            sl->ph_ev.push_back(e);
        }
    auto mark = [&](int k) {
        if (p->profile) HIPCHECK(hipEventRecord(sl->ph_ev[k], st));
    };
    mark(0);
    HIPCHECK(hipMemcpyAsync(sl->dev, h, used, hipMemcpyHostToDevice, st));
    const kun::PoolBlock *dblk = (const kun::PoolBlock *)(sl->dev + off_blocks);
    const kun_cs::Seg *dseg = (const kun_cs::Seg *)(sl->dev + off_segs);
    ScatterArgs sa;
    memset(&sa, 0, sizeof sa);
    for (size_t q = 0; q < nin; q++) sa.dst[q] = p->row(p->inputs[q]);
    sa.packed = sl->dev + off_in;
    sa.total = (long)total;
    sa.blocks = dblk;
    sa.nin = (int)nin;
    if (p->esz == 8)
        hipLaunchKernelGGL(pool_scatter<double>, dim3((unsigned)nblocks), dim3(64), 0, st, sa);
    else
        hipLaunchKernelGGL(pool_scatter<float>, dim3((unsigned)nblocks), dim3(64), 0, st, sa);
    HIPCHECK(hipGetLastError());
    mark(1);
    kun::PoolArgs pa{p->d_args, dblk};
    void *args[] = {&pa};
    size_t so = 0;
    for (int ph = 0; ph < m->nphase; ph++) {
        const KunPhaseDesc &pd = m->phases[ph];
        if (pd.kind == KUN_PHASE_TS) {
            HIPCHECK(hipLaunchKernel(pd.kernel, dim3((unsigned)nblocks), dim3(64), args, 0, st));
            mark(ph + 2);
            continue;
        }
        const int nr = p->ncs_rank[ph] * n, ns = p->ncs_scale[ph] * n;
        hipError_t e = kun_cs::launch_rank_segs(m->dtype, dseg + so, nr, maxS, st);
        if (e == hipSuccess) e = kun_cs::launch_scale_segs(m->dtype, dseg + so + nr, ns, st);
        if (e != hipSuccess) throw Error(std::string("cross-sectional kernel: ") + hipGetErrorString(e));
        so += (size_t)nr + ns;
        mark(ph + 2);
    }
    sl->res_off.clear();
    sl->res_total = 0;
    if (!p->rb_bufs.empty()) {
        GatherArgs ga{p->d_rb_src, sl->res_dev, (long)total, dblk, (int)p->rb_bufs.size()};
        if (p->esz == 8)
            hipLaunchKernelGGL(pool_gather<double>, dim3((unsigned)nblocks), dim3(64), 0, st, ga);
        else
            hipLaunchKernelGGL(pool_gather<float>, dim3((unsigned)nblocks), dim3(64), 0, st, ga);
        HIPCHECK(hipGetLastError());
        HIPCHECK(hipEventRecord(sl->computed, st));
        HIPCHECK(hipMemcpyAsync(sl->res_host, sl->res_dev, p->rb_bufs.size() * total * p->esz,
                                hipMemcpyDeviceToHost, st));
        sl->res_total = total;
        for (int j = 0; j < n; j++) sl->res_off[sessions[j]] = offs[j];
    } else {
        HIPCHECK(hipEventRecord(sl->computed, st));
    }
    mark(m->nphase + 2);
    HIPCHECK(hipEventRecord(sl->done, st));
    sl->busy = true;
    sl->cycle = p->next_cycle++;
    for (int j = 0; j < n; j++) {
        PoolSession &ss = p->sess[sessions[j]];
        ss.tick++;
        ss.slot = (int)(sl - p->slots.data());
        sl->sessions.push_back(sessions[j]);
    }
    return sl->cycle;
}

// 1 when cycle `c` has finished (or is unknown, i.e. finished earlier)
int pool_done(StreamPool *p, unsigned long long c, bool wait) {
    for (auto &sl : p->slots)
        if (sl.busy && sl.cycle == c) {
            if (wait) {
                wait_slot(p, sl);
                return 1;
            }
            hipError_t e = hipEventQuery(sl.done);
            if (e == hipErrorNotReady) return 0;
            HIPCHECK(e);
            retire(p, sl);
            return 1;
        }
    return 1;
}

// 1 when cycle `c` has finished its phases (its readback may still run)
int pool_computed(StreamPool *p, unsigned long long c, bool wait) {
    for (auto &sl : p->slots)
        if (sl.busy && sl.cycle == c) {
            hipError_t e = wait ? hipEventSynchronize(sl.computed) : hipEventQuery(sl.computed);
            if (e == hipErrorNotReady) return 0;
            HIPCHECK(e);
            return 1;
        }
    return 1;
}

const void *pool_row_device(StreamPool *p, int sid, size_t b) {
    const PoolSession &ss = live_session(p, sid);
    if (b >= (size_t)p->m->nbuf) throw Error("bad buffer handle");
    return (const char *)p->row((int)b) + ss.s0 * p->esz;
}

void pool_read(StreamPool *p, int sid, size_t b, void *host) {
    PoolSession &ss = live_session(p, sid);
    if (b >= (size_t)p->m->nbuf) throw Error("bad buffer handle");
    if (ss.slot >= 0) wait_slot(p, p->slots[ss.slot]);
    HIPCHECK(hipSetDevice(p->ex->device));
    HIPCHECK(hipMemcpy(host, (const char *)p->row((int)b) + ss.s0 * p->esz, ss.S * p->esz, hipMemcpyDeviceToHost));
}

long pool_ticks(StreamPool *p, int sid) { return live_session(p, sid).tick; }

void pool_set_readback(StreamPool *p, int n, const size_t *bufs) {
    if (n < 0 || (n > 0 && !bufs)) throw Error("bad readback buffer list");
    for (int i = 0; i < n; i++)
        if (bufs[i] >= (size_t)p->m->nbuf) throw Error("bad buffer handle");
    HIPCHECK(hipSetDevice(p->ex->device));
    for (auto &sl : p->slots) {
        wait_slot(p, sl);
        if (sl.res_host) HIPCHECK(hipHostFree(sl.res_host));
        if (sl.res_dev) HIPCHECK(hipFree(sl.res_dev));
        sl.res_host = sl.res_dev = nullptr;
        sl.res_off.clear();
        sl.res_total = 0;
    }
    if (p->d_rb_src) HIPCHECK(hipFree(p->d_rb_src));
    p->d_rb_src = nullptr;
    p->rb_bufs.assign(bufs, bufs + n);
    if (!n) return;
    std::vector<const void *> src(n);
    for (int i = 0; i < n; i++) src[i] = p->row((int)bufs[i]);
    HIPCHECK(hipMalloc((void **)&p->d_rb_src, n * sizeof(void *)));
    HIPCHECK(hipMemcpy(p->d_rb_src, src.data(), n * sizeof(void *), hipMemcpyHostToDevice));
    const size_t bytes = (size_t)n * p->cap * p->esz;
    for (auto &sl : p->slots) {
        HIPCHECK(hipHostMalloc((void **)&sl.res_host, bytes, hipHostMallocDefault));
        HIPCHECK(hipMalloc((void **)&sl.res_dev, bytes));
    }
}

static PoolSlot &cycle_slot(StreamPool *p, unsigned long long c) {
    for (auto &sl : p->slots)
        if (sl.cycle == c && c) return sl;
    throw Error("cycle " + std::to_string(c) + " is not the last cycle of any slot (its results were replaced)");
}

const void *pool_result(StreamPool *p, unsigned long long c, int sid, size_t b) {
    PoolSlot &sl = cycle_slot(p, c);
    size_t q = 0;
    while (q < p->rb_bufs.size() && p->rb_bufs[q] != b) q++;
    if (q == p->rb_bufs.size()) throw Error("buffer is not a readback buffer of the pool");
    auto it = sl.res_off.find(sid);
    if (it == sl.res_off.end()) throw Error("session " + std::to_string(sid) + " was not in cycle " + std::to_string(c));
    wait_slot(p, sl);
    return sl.res_host + ((size_t)q * sl.res_total + it->second) * p->esz;
}

void pool_profile(StreamPool *p, bool on) { p->profile = on; }

int pool_phase_times(StreamPool *p, unsigned long long c, float *ms, int n) {
    PoolSlot &sl = cycle_slot(p, c);
    if (!sl.profiled) throw Error("cycle " + std::to_string(c) + " was not profiled");
    const int k = p->m->nphase + 2;
    if (n < k) throw Error("phase time array needs " + std::to_string(k) + " entries");
    wait_slot(p, sl);
    for (int i = 0; i < k; i++) HIPCHECK(hipEventElapsedTime(&ms[i], sl.ph_ev[i], sl.ph_ev[i + 1]));
    return k;
}

// One session's rows and stage states as a stream context of S stocks
// would hold them (WireHost), copied out of / into the pool's [..][cap]
// arrays with 2-D copies over the session's lanes.
static void session_pull(StreamPool *p, const PoolSession &ss, WireHost &h) {
    const KunModuleDesc *m = p->m;
    const size_t S = ss.S, esz = p->esz, cap = p->cap;
    h.rows.assign(m->nbuf, std::vector<char>(S * esz));
    HIPCHECK(hipSetDevice(p->ex->device));
    for (int b = 0; b < m->nbuf; b++)
        HIPCHECK(hipMemcpy(h.rows[b].data(), (const char *)p->row(b) + ss.s0 * esz, S * esz, hipMemcpyDeviceToHost));
    h.state.assign(m->nstage, std::vector<char>());
    for (int si = 0; si < m->nstage; si++) {
        const KunStageDesc &sd = m->stages[si];
        if (sd.kind != KUN_STAGE_TS || !sd.state_bytes) continue;
        const size_t nt = (size_t)m->stage_tslots[si], ni = ((size_t)sd.state_bytes - nt * esz) / 4;
        h.state[si].resize((size_t)sd.state_bytes * S);
        const char *base = (const char *)p->state[si];
        if (nt)
            HIPCHECK(hipMemcpy2D(h.state[si].data(), S * esz, base + ss.s0 * esz, cap * esz, S * esz, nt,
                                 hipMemcpyDeviceToHost));
        if (ni)
            HIPCHECK(hipMemcpy2D(h.state[si].data() + nt * S * esz, S * 4, base + nt * cap * esz + ss.s0 * 4, cap * 4,
                                 S * 4, ni, hipMemcpyDeviceToHost));
    }
}

static void session_push(StreamPool *p, const PoolSession &ss, const WireHost &h) {
    const KunModuleDesc *m = p->m;
    const size_t S = ss.S, esz = p->esz, cap = p->cap;
    HIPCHECK(hipSetDevice(p->ex->device));
    for (int b = 0; b < m->nbuf; b++)
        HIPCHECK(hipMemcpy((char *)p->row(b) + ss.s0 * esz, h.rows[b].data(), S * esz, hipMemcpyHostToDevice));
    for (int si = 0; si < m->nstage; si++) {
        const KunStageDesc &sd = m->stages[si];
        if (sd.kind != KUN_STAGE_TS || !sd.state_bytes) continue;
        const size_t nt = (size_t)m->stage_tslots[si], ni = ((size_t)sd.state_bytes - nt * esz) / 4;
        char *base = (char *)p->state[si];
        if (nt)
            HIPCHECK(hipMemcpy2D(base + ss.s0 * esz, cap * esz, h.state[si].data(), S * esz, S * esz, nt,
                                 hipMemcpyHostToDevice));
        if (ni)
            HIPCHECK(hipMemcpy2D(base + nt * cap * esz + ss.s0 * 4, cap * 4, h.state[si].data() + nt * S * esz, S * 4,
                                 S * 4, ni, hipMemcpyHostToDevice));
    }
}

// A session of S stocks resumed from a serialised stream state (the
// reference's byte layout); -1 when the pool has no room. A bad blob
// releases the range again and throws the reference's messages.
int pool_admit_state(StreamPool *p, size_t S, const char *state, size_t len) {
    if (!p->m->stage_tslots) throw Error("the module carries no stream-state table");
    int sid = pool_admit(p, S);
    if (sid < 0) return sid;
    PoolSession &ss = p->sess[sid];
    try {
        WireHost h;
        session_pull(p, ss, h);
        wire_deserialize(WireView{p->m, S, 0, 0, &ss.pos0}, h, state, state + len);
        session_push(p, ss, h);
    } catch (...) {
        pool_release(p, sid);
        throw;
    }
    ss.tick = 0;
    ss.tick0 = 0;
    return sid;
}

// The session's state in the reference's byte layout (waits for its cycle)
std::string pool_serialize(StreamPool *p, int sid) {
    PoolSession &ss = live_session(p, sid);
    if (!p->m->stage_tslots) throw Error("the module carries no stream-state table");
    if (ss.slot >= 0) wait_slot(p, p->slots[ss.slot]);
    WireHost h;
    session_pull(p, ss, h);
    std::string out;
    wire_serialize(WireView{p->m, ss.S, ss.tick, ss.tick0, &ss.pos0}, h, out);
    return out;
}

} // namespace kunrt
