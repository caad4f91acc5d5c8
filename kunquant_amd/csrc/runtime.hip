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
    long nl = env_long("KUN_STAGE_LANES", 6);
    nl = nl < 0 ? 0 : (nl > 16 ? 16 : nl);
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
        auto &acc = stage_ms[std::make_tuple(p.module, p.serial, p.stage, p.kern)];
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

static std::mutex g_serial_mu;
static uint64_t g_next_serial = 0;
static std::unordered_map<const KunModuleDesc *, uint64_t> g_module_serial;

std::shared_ptr<Library> Library::load(const char *path) {
    void *h = dlopen(path, RTLD_NOW | RTLD_LOCAL);
    if (!h) {
        fprintf(stderr, "dlopen failed: %s %s\n", path, dlerror());
        return nullptr;
    }
    auto lib = std::make_shared<Library>();
    lib->handle = h;
    lib->path = path;
    std::lock_guard<std::mutex> g(g_serial_mu);
    lib->serial = ++g_next_serial;
    return lib;
}

const KunModuleDesc *Library::get_module(const char *name) {
    auto *m = (const KunModuleDesc *)dlsym(handle, module_symbol(name).c_str());
    if (m && m->abi != KUN_MODULE_ABI_VERSION) return nullptr;
    if (m) {
        std::lock_guard<std::mutex> g(g_serial_mu);
        g_module_serial[m] = serial;
    }
    return m;
}

uint64_t module_serial(const KunModuleDesc *m) {
    std::lock_guard<std::mutex> g(g_serial_mu);
    auto it = g_module_serial.find(m);
    return it == g_module_serial.end() ? 0 : it->second;
}

// topological order of the stage DAG (cached per descriptor and load serial)
static std::mutex g_order_mu;
static std::map<std::pair<const KunModuleDesc *, uint64_t>, std::vector<int>> g_orders;

const std::vector<int> &stage_order(const KunModuleDesc *m) {
    const auto key = std::make_pair(m, module_serial(m));
    std::lock_guard<std::mutex> g(g_order_mu);
    auto it = g_orders.find(key);
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
    // drop the orders of descriptors that earlier loads placed at this address
    for (auto o = g_orders.begin(); o != g_orders.end();)
        o = (o->first.first == m) ? g_orders.erase(o) : std::next(o);
    return g_orders.emplace(key, std::move(order)).first->second;
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
    p.min_workgroups = env_long("KUN_MIN_KERNEL_WORKGROUPS", p.min_workgroups);
    return p;
}

static size_t align256(size_t x) { return (x + 255) & ~(size_t)255; }
void plan_nodes(const KunModuleDesc *m, size_t S, size_t L, int lanes, BatchLayout &lay);

// Launch table of one stage kernel. Rows of a PAR component are split into
// chunks so that every workgroup gets about the same estimated work
// (comp_cost x rows, halo rows included): the kernel's total work divided by
// `target_workgroups`, each chunk at least `halo_factor` times the halo it
// recomputes and `min_rows` long. Entries run heaviest workgroup first.
static void plan_kernel(const KunKernelDesc &kd, size_t S, size_t L, const ChunkPolicy &pol,
                        BatchLayout::Kern &k, size_t &off) {
    const long nsb = ((long)S + 63) / 64;
    const int nc = kd.ncomp;
    std::vector<long> nch(nc, 1), rows(nc, (long)L > 0 ? (long)L : 1);
    if (kd.kind == KUN_KERNEL_PAR && L > 0) {
        double total = 0;
        for (int z = 0; z < nc; z++) total += (double)kd.comp_cost[z] * (double)L * (double)nsb;
        const double per_wg = std::max(1.0, total / (double)std::max(1L, pol.target_workgroups));
        for (int z = 0; z < nc; z++) {
            const long h = kd.comp_halo[z];
            const long rows_min = std::max(std::max(1L, pol.min_rows), pol.halo_factor * h);
            long lc = (long)(per_wg / (double)kd.comp_cost[z]) - h;
            lc = std::min((long)L, std::max(rows_min, lc));
            nch[z] = ((long)L + lc - 1) / lc;
            rows[z] = ((long)L + nch[z] - 1) / nch[z];
            nch[z] = ((long)L + rows[z] - 1) / rows[z];
        }
        long wgs = 0;
        for (int z = 0; z < nc; z++) wgs += nsb * nch[z];
        if (wgs < pol.min_workgroups) {
            const long want = (pol.min_workgroups + nsb * nc - 1) / (nsb * nc);
            for (int z = 0; z < nc; z++) {
                if (nch[z] >= want) continue;
                const long floor_rows = std::max(16L, (long)kd.comp_halo[z]);
                long lc = std::max(floor_rows, ((long)L + want - 1) / want);
                lc = std::min((long)L, lc);
                const long n = ((long)L + lc - 1) / lc;
                if (n <= nch[z]) continue;
                rows[z] = ((long)L + n - 1) / n;
                nch[z] = ((long)L + rows[z] - 1) / rows[z];
            }
        }
    }
    std::vector<int> order(nc);
    for (int z = 0; z < nc; z++) order[z] = z;
    auto wg_work = [&](int z) {
        const long h = kd.kind == KUN_KERNEL_PAR ? kd.comp_halo[z] : 0;
        return (double)kd.comp_cost[z] * (double)(rows[z] + h);
    };
    std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return wg_work(a) > wg_work(b); });
    k.nz = nc;
    k.blocks = 0;
    for (int e = 0; e < nc; e++) {
        const int z = order[e];
        k.zcomp[e] = z;
        k.zfirst[e] = (int)k.blocks;
        k.zchunk[e] = (int)rows[z];
        k.znch[e] = (int)nch[z];
        k.zscr[e] = (long)(off - k.scratch);
        off += align256((size_t)kd.comp_spill[z] * S * (size_t)nch[z]);
        k.blocks += (unsigned long)(nsb * nch[z]);
    }
    k.zfirst[nc] = (int)k.blocks;
}

// Slots of the [rows][stocks] values of a batch run (module temporaries and
// materialised values). Nodes are visited in launch order; a value written
// by node n takes a slot whose previous value was written and read only by
// nodes that finish before n starts: ancestors of n through the DAG's
// edges and through the order of each lane (a lane is one HIP stream; an
// executor with fewer streams folds lanes together, which keeps each
// lane's order). Slots are added when none is free.
static void plan_slots(const KunModuleDesc *m, size_t S, size_t L, size_t base, BatchLayout &lay) {
    const size_t N = lay.nodes.size();
    const size_t W = (N + 63) / 64;
    std::vector<uint64_t> anc(N * W, 0);
    std::map<int, int> last_on_lane;
    for (size_t n = 0; n < N; n++) {
        std::vector<int> pred = lay.nodes[n].deps;
        auto it = last_on_lane.find(lay.nodes[n].lane);
        if (it != last_on_lane.end()) pred.push_back(it->second);
        last_on_lane[lay.nodes[n].lane] = (int)n;
        for (int d : pred) {
            for (size_t w = 0; w < W; w++) anc[n * W + w] |= anc[(size_t)d * W + w];
            anc[n * W + (size_t)d / 64] |= 1ull << (d % 64);
        }
    }
    auto is_anc = [&](size_t n, int a) { return (anc[n * W + (size_t)a / 64] >> (a % 64)) & 1; };
    // value ids: module buffer b, or nbuf + first_mat[stage] + m
    std::vector<int> first_mat(m->nstage + 1, 0);
    for (int i = 0; i < m->nstage; i++)
        first_mat[i + 1] = first_mat[i] + (m->stages[i].kind == KUN_STAGE_TS ? m->stages[i].nmat : 0);
    const int nval = m->nbuf + first_mat[m->nstage];
    std::vector<std::vector<int>> touch(nval); // nodes writing or reading each value
    std::vector<std::vector<int>> writes(N);
    auto is_temp = [&](int b) { return m->bufs[b].kind == KUN_BUF_TEMP; };
    for (size_t n = 0; n < N; n++) {
        const auto &nd = lay.nodes[n];
        const KunStageDesc &sd = m->stages[nd.stage];
        auto rd = [&](int v) { touch[v].push_back((int)n); };
        auto wr = [&](int v) {
            touch[v].push_back((int)n);
            writes[n].push_back(v);
        };
        if (sd.kind == KUN_STAGE_TS) {
            const KunKernelDesc &kd = sd.kerns[nd.kern];
            for (int r = 0; r < kd.nreads; r++)
                if (is_temp(kd.reads[r])) rd(kd.reads[r]);
            for (int w = 0; w < kd.nwrites; w++)
                if (is_temp(kd.writes[w])) wr(kd.writes[w]);
            const int mb = m->nbuf + first_mat[nd.stage];
            for (int r = 0; r < kd.nmat_reads; r++) rd(mb + kd.mat_reads[r]);
            for (int w = 0; w < kd.nmat_writes; w++) wr(mb + kd.mat_writes[w]);
        } else {
            if (is_temp(sd.in[0])) rd(sd.in[0]);
            if (is_temp(sd.out[0])) wr(sd.out[0]);
        }
    }
    std::vector<int> slot_of(nval, -1), occupant;
    auto place = [&](size_t n, int v) {
        if (slot_of[v] >= 0) return;
        for (size_t s = 0; s < occupant.size(); s++) {
            bool free_ = true;
            for (int t : touch[occupant[s]])
                if (!is_anc(n, t)) {
                    free_ = false;
                    break;
                }
            if (free_) {
                occupant[s] = v;
                slot_of[v] = (int)s;
                return;
            }
        }
        slot_of[v] = (int)occupant.size();
        occupant.push_back(v);
    };
    for (size_t n = 0; n < N; n++)
        for (int v : writes[n]) place(n, v);
    // values nothing writes (unused temporaries) keep slots of their own
    for (int v = 0; v < nval; v++)
        if (slot_of[v] < 0 && (v >= m->nbuf || is_temp(v))) {
            slot_of[v] = (int)occupant.size();
            occupant.push_back(v);
        }
    lay.slot_bytes = align256(std::max<size_t>(1, S * L) * elem_size(m));
    lay.nslots = (int)occupant.size();
    auto off = [&](int v) { return base + (size_t)slot_of[v] * lay.slot_bytes; };
    lay.buf_off.assign(m->nbuf, -1);
    for (int b = 0; b < m->nbuf; b++)
        if (is_temp(b)) lay.buf_off[b] = (long)off(b);
    lay.mat_off.assign(m->nstage, {});
    for (int i = 0; i < m->nstage; i++)
        for (int j = first_mat[i]; j < first_mat[i + 1]; j++) lay.mat_off[i].push_back(off(m->nbuf + j));
    lay.bytes = base + (size_t)lay.nslots * lay.slot_bytes;
}

BatchLayout batch_layout(const KunModuleDesc *m, size_t S, size_t L, const ChunkPolicy &pol, int lanes) {
    BatchLayout lay;
    lay.kern.resize(m->nstage);
    lay.cs_scratch.assign(m->nstage, 0);
    size_t off = 0;
    for (int i = 0; i < m->nstage; i++) {
        const KunStageDesc &sd = m->stages[i];
        if (sd.kind == KUN_STAGE_SCALE) {
            lay.cs_scratch[i] = off;
            off += align256(kun_cs::scale_scratch_bytes(m->dtype, (long)L));
        }
        if (sd.kind != KUN_STAGE_TS) continue;
        if (sd.nin + sd.nout + sd.nmat > 256) throw Error("stage with too many buffers");
        for (int j = 0; j < sd.nkern; j++) {
            if (sd.kerns[j].ncomp > KUN_MAX_COMPS) throw Error("stage kernel with too many components");
            BatchLayout::Kern k;
            k.scratch = off;
            plan_kernel(sd.kerns[j], S, L, pol, k, off);
            lay.kern[i].push_back(k);
        }
    }
    plan_nodes(m, S, L, lanes, lay);
    plan_slots(m, S, L, align256(off), lay);
    return lay;
}

// Kernels of kernel `kj` of stage `si` (cross-sectional stages: kj = 0)
// into `sink` (a stream, or a graph under construction). prof: per-kernel
// event timing on stream `st` (stream sinks only).
static void enqueue_node(const KunModuleDesc *m, int si, int kj, const std::vector<void *> &dev,
                         const BatchLayout &lay, char *arena, size_t S, size_t T, size_t start, size_t L,
                         kun_cs::Sink &sink, hipStream_t st, Executor *prof) {
    const KunStageDesc &sd = m->stages[si];
    Executor::Pending ev{m, prof ? module_serial(m) : 0, si, kj, nullptr, nullptr};
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
        for (int j = 0; j < sd.nmat; j++) a.bufs[n++] = arena + lay.mat_off[si][j];
        a.S = (long)S;
        a.T = (long)T;
        a.start = (long)start;
        a.L = (long)L;
        a.blk = m->blocking;
        a.scratch = arena + kl.scratch;
        a.nz = kl.nz;
        memcpy(a.zcomp, kl.zcomp, sizeof a.zcomp);
        memcpy(a.zfirst, kl.zfirst, sizeof a.zfirst);
        memcpy(a.zchunk, kl.zchunk, sizeof a.zchunk);
        memcpy(a.znch, kl.znch, sizeof a.znch);
        memcpy(a.zscr, kl.zscr, sizeof a.zscr);
        void *args[] = {&a};
        if (kl.blocks > 0x7fffffffUL) throw Error("batch too large for one launch");
        hipError_t e = sink.launch(kd.kernel, dim3((unsigned)kl.blocks), dim3(sd.block), args);
        if (e != hipSuccess) throw Error(std::string("stage kernel launch: ") + hipGetErrorString(e));
    } else if (sd.kind == KUN_STAGE_GENERIC) {
        kun::CsArgs a;
        memset(&a, 0, sizeof a);
        int n = 0;
        auto add = [&](int b) {
            kun_cs::View v = cs_view(m, b, dev[b], T, start, L);
            a.bufs[n] = v.ptr;
            a.lay[n] = v.layout;
            a.rows[n] = v.rows;
            a.row0[n] = v.row0;
            n++;
        };
        if (sd.nin + sd.nout > KUN_MAX_CS_BUFS) throw Error("too many buffers in a cross-sectional stage");
        for (int j = 0; j < sd.nin; j++) add(sd.in[j]);
        for (int j = 0; j < sd.nout; j++) add(sd.out[j]);
        a.S = (long)S;
        a.L = (long)L;
        a.blk = m->blocking;
        void *args[] = {&a};
        hipError_t e = sink.launch(sd.stream_kernel, dim3((unsigned)((L + 63) / 64)), dim3(64), args);
        if (e != hipSuccess) throw Error(std::string("cross-sectional kernel launch: ") + hipGetErrorString(e));
    } else {
        int ib = sd.in[0], ob = sd.out[0];
        kun_cs::View vi = cs_view(m, ib, dev[ib], T, start, L);
        kun_cs::View vo = cs_view(m, ob, dev[ob], T, start, L);
        if (sd.kind == KUN_STAGE_RANK && (long)S > kun_cs::max_rank_stocks(m->dtype))
            throw Error("Rank over more than " + std::to_string(kun_cs::max_rank_stocks(m->dtype)) +
                        " stocks is not supported");
        hipError_t e = sd.kind == KUN_STAGE_RANK
                           ? kun_cs::launch_rank(m->dtype, vi, vo, (long)S, (long)L, m->blocking, sink)
                           : kun_cs::launch_scale(m->dtype, vi, vo, (long)S, (long)L, m->blocking, sink,
                                                  arena + lay.cs_scratch[si]);
        if (e != hipSuccess) throw Error(std::string("cross-sectional kernel: ") + hipGetErrorString(e));
    }
    if (prof) {
        HIPCHECK(hipEventRecord(ev.stop, st));
        prof->pending.push_back(ev);
    }
}

static void enqueue_node(const KunModuleDesc *m, int si, int kj, const std::vector<void *> &dev,
                         const BatchLayout &lay, char *arena, size_t S, size_t T, size_t start, size_t L,
                         hipStream_t st, Executor *prof) {
    kun_cs::StreamSink sink(st);
    enqueue_node(m, si, kj, dev, lay, arena, S, T, start, L, sink, st, prof);
}

// Launch DAG of a batch run: one node per stage kernel (one per
// cross-sectional stage), in a topological order. A node depends on the
// node writing each buffer it reads and on the kernels of its stage whose
// materialised values it reads. Nodes are list-scheduled over `lanes`
// streams with rough duration estimates (earliest estimated start first).
void plan_nodes(const KunModuleDesc *m, size_t S, size_t L, int lanes, BatchLayout &lay) {
    lay.nodes.clear();
    std::vector<int> writer(m->nbuf, -1);
    std::vector<std::vector<int>> node_of(m->nstage);
    const double nsb = (double)((S + 63) / 64);
    for (int si : stage_order(m)) {
        const KunStageDesc &sd = m->stages[si];
        const int nk = sd.kind == KUN_STAGE_TS ? sd.nkern : 1;
        for (int kj = 0; kj < nk; kj++) {
            BatchLayout::Node nd;
            nd.stage = si;
            nd.kern = kj;
            const int id = (int)lay.nodes.size();
            if (sd.kind == KUN_STAGE_TS) {
                const KunKernelDesc &kd = sd.kerns[kj];
                for (int r = 0; r < kd.nreads; r++)
                    if (writer[kd.reads[r]] >= 0) nd.deps.push_back(writer[kd.reads[r]]);
                for (int r = 0; r < kd.nkdeps; r++) nd.deps.push_back(node_of[si][kd.kdeps[r]]);
                for (int w = 0; w < kd.nwrites; w++) writer[kd.writes[w]] = id;
                // microseconds: PAR ~ total work at ~40% of the VALU peak;
                // SEQ ~ one dependent walk over the rows
                double work = 0, seq = 0;
                for (int z = 0; z < kd.ncomp; z++) {
                    work += (double)kd.comp_cost[z] * ((double)L + (double)kd.comp_halo[z]) * nsb;
                    seq = std::max(seq, (double)kd.comp_cost[z]);
                }
                nd.est = kd.kind == KUN_KERNEL_PAR ? 5.0 + work * 4.0 / (1024.0 * 2400.0 * 0.4)
                                                   : 20.0 + (double)L * seq * 8.0 / 2400.0;
            } else {
                if (writer[sd.in[0]] >= 0) nd.deps.push_back(writer[sd.in[0]]);
                writer[sd.out[0]] = id;
                nd.est = 10.0 + (double)L * (sd.kind == KUN_STAGE_RANK ? 0.18 : 0.12) * (double)S / 8192.0;
            }
            std::sort(nd.deps.begin(), nd.deps.end());
            nd.deps.erase(std::unique(nd.deps.begin(), nd.deps.end()), nd.deps.end());
            node_of[si].push_back(id);
            lay.nodes.push_back(nd);
        }
    }
    // list scheduling in topological order
    std::vector<double> lane_free(std::max(1, lanes), 0.0), finish(lay.nodes.size(), 0.0);
    for (size_t n = 0; n < lay.nodes.size(); n++) {
        BatchLayout::Node &nd = lay.nodes[n];
        double ready = 0;
        int last_dep = -1;
        for (int d : nd.deps)
            if (finish[d] >= ready) {
                ready = finish[d];
                last_dep = d;
            }
        int best = 0;
        double best_start = 1e300;
        for (int l = 0; l < (int)lane_free.size(); l++) {
            double st = std::max(lane_free[l], ready);
            bool pref = last_dep >= 0 && lay.nodes[last_dep].lane == l;
            if (st < best_start - 1e-9 || (st <= best_start + 1e-9 && pref)) {
                best_start = st;
                best = l;
            }
        }
        nd.lane = best;
        finish[n] = best_start + nd.est;
        lane_free[best] = finish[n];
    }
}

// Enqueue every kernel of a batch module; dev[b] are device pointers, arena
// the scratch described by `lay`. Nodes run on the executor's lanes (HIP
// streams) as planned by plan_nodes; a node waits through events for its
// dependencies on other lanes, every lane waits for `st` at the start and
// `st` waits for every lane at the end.
void enqueue_batch_serial(const KunModuleDesc *m, const std::vector<void *> &dev, const BatchLayout &lay,
                          char *arena, size_t S, size_t T, size_t start, size_t L, hipStream_t st, Executor *prof) {
    if (L == 0) return;
    for (const auto &nd : lay.nodes) enqueue_node(m, nd.stage, nd.kern, dev, lay, arena, S, T, start, L, st, prof);
}

void enqueue_batch(Executor &ex, const KunModuleDesc *m, const std::vector<void *> &dev, const BatchLayout &lay,
                   char *arena, size_t S, size_t T, size_t start, size_t L, hipStream_t st) {
    if (L == 0) return;
    Executor *prof = ex.profiling ? &ex : nullptr;
    const int nl = (int)ex.lanes.size();
    if (nl == 0) {
        enqueue_batch_serial(m, dev, lay, arena, S, T, start, L, st, prof);
        return;
    }
    if (lay.sync_lanes != nl) {
        std::vector<int> lane(lay.nodes.size());
        std::vector<std::vector<int>> deps(lay.nodes.size());
        for (size_t n = 0; n < lay.nodes.size(); n++) {
            lane[n] = lay.nodes[n].lane % nl;
            deps[n] = lay.nodes[n].deps;
        }
        BatchLayout &w = const_cast<BatchLayout &>(lay); // a cache of the layout's own plan
        w.sync = plan_lane_sync(lane, deps, nl);
        w.sync_lanes = nl;
    }
    ex.ensure_stage_events((int)lay.nodes.size());
    HIPCHECK(hipEventRecord(ex.fork, st));
    for (int l = 0; l < nl; l++) HIPCHECK(hipStreamWaitEvent(ex.lanes[l], ex.fork, 0));
    for (size_t n = 0; n < lay.nodes.size(); n++) {
        const auto &nd = lay.nodes[n];
        hipStream_t ls = ex.lanes[nd.lane % nl];
        for (int d : lay.sync.waits[n]) HIPCHECK(hipStreamWaitEvent(ls, ex.stage_done[d], 0));
        enqueue_node(m, nd.stage, nd.kern, dev, lay, arena, S, T, start, L, ls, prof);
        if (lay.sync.record[n]) HIPCHECK(hipEventRecord(ex.stage_done[n], ls));
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
    struct Release { // pooled inputs and outputs (temporaries live in the arena)
        Executor &ex;
        const KunModuleDesc *m;
        std::vector<void *> &dev;
        std::vector<size_t> &bytes;
        ~Release() {
            hipStreamSynchronize(ex.stream);
            for (size_t i = 0; i < dev.size(); i++)
                if (m->bufs[i].kind != KUN_BUF_TEMP) ex.pool.put(dev[i], bytes[i]);
        }
    } release{ex, m, dev, bytes};
    for (int b = 0; b < m->nbuf; b++) {
        const KunBufferDesc &d = m->bufs[b];
        if (d.kind != KUN_BUF_TEMP && !bufs.count(d.name))
            throw Error(std::string("Buffer name not found: ") + d.name);
    }
    BatchLayout lay = batch_layout(m, S, L, ex.policy_for(m, S, L), (int)ex.lanes.size());
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
    for (int b = 0; b < m->nbuf; b++) {
        const KunBufferDesc &d = m->bufs[b];
        if (d.kind == KUN_BUF_TEMP) {
            dev[b] = arena.p + lay.buf_off[b];
            continue;
        }
        bytes[b] = buffer_elems(m, b, S, T, L) * esz;
        dev[b] = ex.pool.get(bytes[b]);
        if (d.kind == KUN_BUF_INPUT)
            HIPCHECK(hipMemcpyAsync(dev[b], bufs.at(d.name), bytes[b], hipMemcpyHostToDevice, ex.stream));
    }
    enqueue_batch(ex, m, dev, lay, arena.p, S, T, start, L, ex.stream);
    for (int b = 0; b < m->nbuf; b++) {
        const KunBufferDesc &d = m->bufs[b];
        if (d.kind == KUN_BUF_OUTPUT)
            HIPCHECK(hipMemcpyAsync(bufs.at(d.name), dev[b], bytes[b], hipMemcpyDeviceToHost, ex.stream));
    }
    HIPCHECK(hipStreamSynchronize(ex.stream));
}

// Device-buffer run: temporaries from the pool (kept for reuse per shape).
ChunkPolicy Executor::policy_for(const KunModuleDesc *m, size_t S, size_t L) const {
    auto it = tuned.find(std::make_tuple((const void *)m, module_serial(m), S, L));
    return it == tuned.end() ? chunking : it->second;
}

DeviceTemps make_temps(Executor &ex, const KunModuleDesc *m, size_t S, size_t T, size_t L) {
    (void)T;
    DeviceTemps t;
    t.lay = batch_layout(m, S, L, ex.policy_for(m, S, L), (int)ex.lanes.size());
    t.bufs.assign(m->nbuf, nullptr);
    HIPCHECK(hipEventCreateWithFlags(&t.done, hipEventDisableTiming));
    if (t.lay.bytes) {
        hipError_t e = hipMalloc((void **)&t.arena, t.lay.bytes);
        if (e != hipSuccess) {
            hipEventDestroy(t.done);
            throw Error(std::string("hipMalloc of a batch arena failed: ") + hipGetErrorString(e));
        }
    }
    for (int b = 0; b < m->nbuf; b++)
        if (m->bufs[b].kind == KUN_BUF_TEMP) t.bufs[b] = t.arena + t.lay.buf_off[b];
    return t;
}

void free_device_temps(DeviceTemps &t) {
    if (t.arena) hipFree(t.arena);
    if (t.done) hipEventDestroy(t.done);
    t.bufs.clear();
    t.arena = nullptr;
    t.done = nullptr;
    t.armed = false;
}

// A run on `st` that uses arena `t`: wait until the previous run using it
// (possibly on another stream) has finished with it ...
static void arena_acquire(DeviceTemps &t, hipStream_t st) {
    if (t.armed) HIPCHECK(hipStreamWaitEvent(st, t.done, 0));
}
// ... and mark this run's end on `st` (after its lanes joined into `st`)
static void arena_release(DeviceTemps &t, hipStream_t st) {
    HIPCHECK(hipEventRecord(t.done, st));
    t.armed = true;
}

// temporaries of a device-resident run, cached per (module, load serial,
// shape, job slot); entries of a descriptor from an earlier load at the same
// address are freed first
static DeviceTemps &cached_temps(Executor &ex, const KunModuleDesc *m, size_t S, size_t T, size_t L, int slot) {
    const uint64_t serial = module_serial(m);
    for (auto it = ex.temps.begin(); it != ex.temps.end();) {
        if (std::get<0>(it->first) == (const void *)m && std::get<1>(it->first) != serial) {
            HIPCHECK(hipDeviceSynchronize());
            free_device_temps(it->second);
            it = ex.temps.erase(it);
        } else {
            ++it;
        }
    }
    auto key = std::make_tuple((const void *)m, serial, S, L, slot);
    auto it = ex.temps.find(key);
    if (it == ex.temps.end()) it = ex.temps.emplace(key, make_temps(ex, m, S, T, L)).first;
    return it->second;
}

static std::vector<void *> device_pointers(const KunModuleDesc *m, const NameMap &bufs, const DeviceTemps &temps) {
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
    return dev;
}

void run_graph_device(Executor &ex, const KunModuleDesc *m, const NameMap &bufs, size_t S, size_t T,
                      size_t start, size_t L, hipStream_t st) {
    check_batch_module(m, S, T, start, L);
    std::lock_guard<std::mutex> g(ex.mu);
    HIPCHECK(hipSetDevice(ex.device));
    if (!st) st = ex.stream;
    DeviceTemps &temps = cached_temps(ex, m, S, T, L, 0);
    const std::vector<void *> dev = device_pointers(m, bufs, temps);
    arena_acquire(temps, st);
    enqueue_batch(ex, m, dev, temps.lay, temps.arena, S, T, start, L, st);
    arena_release(temps, st);
}

// Several independent jobs enqueued together: their launch DAGs interleave
// over the lanes (job j shifts its lanes by j), so small jobs share the GPU.
static void enqueue_jobs(Executor &ex, const std::vector<BatchJob> &jobs, const std::vector<const DeviceTemps *> &temps,
                         hipStream_t st, bool profile) {
    const int nl = (int)ex.lanes.size();
    Executor *prof = profile && ex.profiling ? &ex : nullptr;
    std::vector<std::vector<void *>> dev(jobs.size());
    std::vector<size_t> first(jobs.size() + 1, 0);
    for (size_t j = 0; j < jobs.size(); j++) {
        dev[j] = device_pointers(jobs[j].m, *jobs[j].bufs, *temps[j]);
        first[j + 1] = first[j] + temps[j]->lay.nodes.size();
    }
    if (nl == 0) {
        for (size_t j = 0; j < jobs.size(); j++)
            for (const auto &nd : temps[j]->lay.nodes)
                enqueue_node(jobs[j].m, nd.stage, nd.kern, dev[j], temps[j]->lay, temps[j]->arena, jobs[j].S,
                             jobs[j].T, jobs[j].start, jobs[j].L, st, prof);
        return;
    }
    ex.ensure_stage_events((int)first.back());
    HIPCHECK(hipEventRecord(ex.fork, st));
    for (int l = 0; l < nl; l++) HIPCHECK(hipStreamWaitEvent(ex.lanes[l], ex.fork, 0));
    // enqueue order: one node of each job in turn (job j shifted j lanes),
    // items numbered in that order; events are indexed by job and node
    std::vector<std::pair<size_t, size_t>> order; // (job, node)
    std::vector<size_t> next(jobs.size(), 0);
    for (bool more = true; more;) {
        more = false;
        for (size_t j = 0; j < jobs.size(); j++)
            if (next[j] < temps[j]->lay.nodes.size()) {
                order.emplace_back(j, next[j]++);
                more = true;
            }
    }
    std::vector<int> item_of(first.back()), lane(order.size());
    std::vector<std::vector<int>> deps(order.size());
    for (size_t i = 0; i < order.size(); i++) {
        const size_t j = order[i].first, n = order[i].second;
        const auto &nd = temps[j]->lay.nodes[n];
        item_of[first[j] + n] = (int)i;
        lane[i] = (int)((nd.lane + j) % nl);
        for (int d : nd.deps) deps[i].push_back(item_of[first[j] + d]);
    }
    const LaneSync sync = plan_lane_sync(lane, deps, nl);
    for (size_t i = 0; i < order.size(); i++) {
        const size_t j = order[i].first, n = order[i].second;
        const auto &nd = temps[j]->lay.nodes[n];
        hipStream_t ls = ex.lanes[lane[i]];
        for (int d : sync.waits[i])
            HIPCHECK(hipStreamWaitEvent(ls, ex.stage_done[first[order[d].first] + order[d].second], 0));
        enqueue_node(jobs[j].m, nd.stage, nd.kern, dev[j], temps[j]->lay, temps[j]->arena, jobs[j].S, jobs[j].T,
                     jobs[j].start, jobs[j].L, ls, prof);
        if (sync.record[i]) HIPCHECK(hipEventRecord(ex.stage_done[first[j] + n], ls));
    }
    for (int l = 0; l < nl; l++) {
        HIPCHECK(hipEventRecord(ex.lane_done[l], ex.lanes[l]));
        HIPCHECK(hipStreamWaitEvent(st, ex.lane_done[l], 0));
    }
}

static void check_jobs(const std::vector<BatchJob> &jobs) {
    for (const auto &jb : jobs) check_batch_module(jb.m, jb.S, jb.T, jb.start, jb.L);
}

void run_jobs_device(Executor &ex, const std::vector<BatchJob> &jobs, hipStream_t st) {
    check_jobs(jobs);
    std::lock_guard<std::mutex> g(ex.mu);
    HIPCHECK(hipSetDevice(ex.device));
    if (!st) st = ex.stream;
    std::vector<DeviceTemps *> own;
    for (size_t j = 0; j < jobs.size(); j++)
        own.push_back(&cached_temps(ex, jobs[j].m, jobs[j].S, jobs[j].T, jobs[j].L, (int)j + 1));
    std::vector<const DeviceTemps *> temps(own.begin(), own.end());
    for (auto *t : own) arena_acquire(*t, st);
    enqueue_jobs(ex, jobs, temps, st, true);
    for (auto *t : own) arena_release(*t, st);
}

// ---------------------------------------------------------------------------
// A batch of jobs as HIP graphs, built node by node (hipGraphAddKernelNode)
// from the jobs' launch DAGs rather than captured from streams: no events
// inside the graphs, no stream capture. Each DAG node contributes its kernels, the first
// depending on the last kernel of each of the node's dependencies and of the
// node before it on the same lane, later ones on the kernel before them;
// jobs are independent subgraphs. The lane edges matter: the arena's memory
// plan (plan_slots) reuses a slot once its previous users precede the new
// writer through DAG edges or through the order of a lane, as on the
// executor's streams.

namespace {
struct GraphSink : kun_cs::Sink {
    GraphBuilder &gb;
    std::vector<void *> deps; // dependencies of the next kernel
    void *last = nullptr;     // last kernel added
    explicit GraphSink(GraphBuilder &g) : gb(g) {}
    hipError_t launch(const void *func, dim3 grid, dim3 block, void **args) override {
        void *node = nullptr;
        hipError_t e = gb.add_kernel(func, grid, block, args, deps, &node);
        if (e == hipSuccess) {
            deps.assign(1, node);
            last = node;
        }
        return e;
    }
};

struct HipGraphBuilder : GraphBuilder {
    hipGraph_t g;
    explicit HipGraphBuilder(hipGraph_t graph) : g(graph) {}
    hipError_t add_kernel(const void *func, dim3 grid, dim3 block, void **args, const std::vector<void *> &deps,
                          void **node) override {
        hipKernelNodeParams p;
        memset(&p, 0, sizeof p);
        p.func = const_cast<void *>(func);
        p.gridDim = grid;
        p.blockDim = block;
        p.kernelParams = args;
        p.sharedMemBytes = 0;
        p.extra = nullptr;
        hipGraphNode_t n = nullptr;
        hipError_t e = hipGraphAddKernelNode(&n, g, (const hipGraphNode_t *)deps.data(), deps.size(), &p);
        *node = n;
        return e;
    }
    hipError_t add_empty(const std::vector<void *> &deps, void **node) override {
        hipGraphNode_t n = nullptr;
        hipError_t e = hipGraphAddEmptyNode(&n, g, (const hipGraphNode_t *)deps.data(), deps.size());
        *node = n;
        return e;
    }
};
} // namespace

void build_jobs_graph(GraphBuilder &gb, const std::vector<BatchJob> &jobs, const std::vector<const DeviceTemps *> &temps) {
    for (size_t j = 0; j < jobs.size(); j++) {
        const BatchJob &jb = jobs[j];
        if (jb.L == 0) continue;
        const DeviceTemps &t = *temps[j];
        const std::vector<void *> dev = device_pointers(jb.m, *jb.bufs, t);
        std::vector<void *> out(t.lay.nodes.size(), nullptr);
        std::map<int, int> last_on_lane;
        for (size_t n = 0; n < t.lay.nodes.size(); n++) {
            const auto &nd = t.lay.nodes[n];
            GraphSink sink(gb);
            std::vector<int> pred = nd.deps;
            auto ln = last_on_lane.find(nd.lane);
            if (ln != last_on_lane.end() && std::find(pred.begin(), pred.end(), ln->second) == pred.end())
                pred.push_back(ln->second);
            last_on_lane[nd.lane] = (int)n;
            for (int d : pred) {
                if (d < 0 || (size_t)d >= n || !out[d]) throw Error("launch DAG out of topological order");
                sink.deps.push_back(out[d]);
            }
            gb.tag((int)j, nd.stage, nd.kern);
            enqueue_node(jb.m, nd.stage, nd.kern, dev, t.lay, t.arena, jb.S, jb.T, jb.start, jb.L, sink, nullptr,
                         nullptr);
            if (!sink.last) {
                hipError_t e = gb.add_empty(sink.deps, &sink.last);
                if (e != hipSuccess) throw Error(std::string("graph node: ") + hipGetErrorString(e));
            }
            out[n] = sink.last;
        }
    }
}

CapturedBatch *capture_jobs(std::shared_ptr<Executor> exp, const std::vector<BatchJob> &jobs) {
    check_jobs(jobs);
    Executor &ex = *exp;
    std::lock_guard<std::mutex> g(ex.mu);
    HIPCHECK(hipSetDevice(ex.device));
    std::unique_ptr<CapturedBatch> cb(new CapturedBatch());
    cb->device = ex.device;
    cb->ex = exp;
    cb->temps.reserve(jobs.size());
    for (const auto &jb : jobs) cb->temps.push_back(make_temps(ex, jb.m, jb.S, jb.T, jb.L));
    // groups of about equal work (stocks x rows x stages, largest first onto
    // the lightest group), as many as the executor has lanes
    const size_t K = std::max<size_t>(1, std::min(ex.lanes.size(), jobs.size()));
    std::vector<size_t> order(jobs.size());
    for (size_t j = 0; j < jobs.size(); j++) order[j] = j;
    auto work = [&](size_t j) { return (double)jobs[j].S * (double)jobs[j].L * (double)jobs[j].m->nstage; };
    std::stable_sort(order.begin(), order.end(), [&](size_t a, size_t b) { return work(a) > work(b); });
    std::vector<std::vector<size_t>> groups(K);
    std::vector<double> load(K, 0.0);
    for (size_t j : order) {
        if (jobs[j].L == 0) continue;
        const size_t k = (size_t)(std::min_element(load.begin(), load.end()) - load.begin());
        groups[k].push_back(j);
        load[k] += work(j);
    }
    HIPCHECK(hipEventCreateWithFlags(&cb->done, hipEventDisableTiming));
    HIPCHECK(hipEventCreateWithFlags(&cb->fork, hipEventDisableTiming));
    for (const auto &grp : groups) {
        if (grp.empty()) continue;
        std::vector<BatchJob> gj;
        std::vector<const DeviceTemps *> gt;
        for (size_t j : grp) {
            gj.push_back(jobs[j]);
            gt.push_back(&cb->temps[j]);
        }
        hipGraph_t graph = nullptr;
        HIPCHECK(hipGraphCreate(&graph, 0));
        cb->graphs.push_back(graph);
        HipGraphBuilder gb(graph);
        build_jobs_graph(gb, gj, gt);
        hipGraphExec_t exec = nullptr;
        HIPCHECK(hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0));
        cb->execs.push_back(exec);
        hipStream_t s;
        HIPCHECK(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
        cb->streams.push_back(s);
        hipEvent_t e;
        HIPCHECK(hipEventCreateWithFlags(&e, hipEventDisableTiming));
        cb->joined.push_back(e);
    }
    return cb.release();
}

void launch_captured(CapturedBatch &cb, hipStream_t st) {
    HIPCHECK(hipSetDevice(cb.device));
    if (!st) st = cb.ex->stream;
    // replays on other streams wait for the previous replay (shared temporaries)
    if (cb.armed) HIPCHECK(hipStreamWaitEvent(st, cb.done, 0));
    HIPCHECK(hipEventRecord(cb.fork, st));
    for (size_t k = 0; k < cb.execs.size(); k++) {
        HIPCHECK(hipStreamWaitEvent(cb.streams[k], cb.fork, 0));
        HIPCHECK(hipGraphLaunch(cb.execs[k], cb.streams[k]));
        HIPCHECK(hipEventRecord(cb.joined[k], cb.streams[k]));
    }
    for (auto e : cb.joined) HIPCHECK(hipStreamWaitEvent(st, e, 0));
    HIPCHECK(hipEventRecord(cb.done, st));
    cb.armed = true;
}

CapturedBatch::~CapturedBatch() {
    if (done) hipEventSynchronize(done);
    for (auto s : streams) hipStreamSynchronize(s);
    for (auto e : execs) hipGraphExecDestroy(e);
    for (auto g : graphs) hipGraphDestroy(g);
    for (auto s : streams) hipStreamDestroy(s);
    for (auto e : joined) hipEventDestroy(e);
    if (fork) hipEventDestroy(fork);
    if (done) hipEventDestroy(done);
    for (auto &t : temps) free_device_temps(t);
}

// Host-only description of the graph capture_jobs would build (no device
// calls): one line per node, "node job stage kernel grid block deps...".
namespace {
struct RecordingBuilder : GraphBuilder {
    std::string text;
    int count = 0, job = 0, stage = 0, kern = 0;
    std::vector<std::unique_ptr<int>> ids;
    void tag(int j, int s, int k) override {
        job = j;
        stage = s;
        kern = k;
    }
    void *record(const char *what, dim3 grid, dim3 block, const std::vector<void *> &deps) {
        ids.emplace_back(new int(count++));
        void *h = ids.back().get();
        text += std::to_string(*(int *)h) + " " + what + " " + std::to_string(job) + " " + std::to_string(stage) +
                " " + std::to_string(kern) + " " + std::to_string(grid.x) + " " + std::to_string(block.x) + " :";
        for (void *d : deps) text += " " + std::to_string(*(int *)d);
        text += "\n";
        return h;
    }
    hipError_t add_kernel(const void *, dim3 grid, dim3 block, void **, const std::vector<void *> &deps,
                          void **node) override {
        *node = record("kernel", grid, block, deps);
        return hipSuccess;
    }
    hipError_t add_empty(const std::vector<void *> &deps, void **node) override {
        *node = record("empty", dim3(0), dim3(0), deps);
        return hipSuccess;
    }
};
} // namespace

std::string describe_jobs_graph(const std::vector<BatchJob> &jobs, int lanes) {
    check_jobs(jobs);
    std::vector<DeviceTemps> own(jobs.size());
    std::vector<const DeviceTemps *> temps;
    for (size_t j = 0; j < jobs.size(); j++) {
        own[j].lay = batch_layout(jobs[j].m, jobs[j].S, jobs[j].L, ChunkPolicy::from_env(), lanes);
        // a placeholder arena base (never dereferenced): the launches see a
        // non-null arena as on the device
        own[j].arena = (char *)(uintptr_t)((j + 1) << 20);
        own[j].bufs.assign(jobs[j].m->nbuf, nullptr);
        for (int b = 0; b < jobs[j].m->nbuf; b++)
            if (jobs[j].m->bufs[b].kind == KUN_BUF_TEMP) own[j].bufs[b] = own[j].arena + own[j].lay.buf_off[b];
        temps.push_back(&own[j]);
    }
    RecordingBuilder rb;
    build_jobs_graph(rb, jobs, temps);
    for (auto &t : own) t.arena = nullptr; // not memory of ours
    // the launch DAGs the graph follows: "dag job node stage kernel lane : deps"
    for (size_t j = 0; j < jobs.size(); j++)
        for (size_t n = 0; n < own[j].lay.nodes.size(); n++) {
            const auto &nd = own[j].lay.nodes[n];
            rb.text += "dag " + std::to_string(j) + " " + std::to_string(n) + " " + std::to_string(nd.stage) + " " +
                       std::to_string(nd.kern) + " " + std::to_string(nd.lane) + " :";
            for (int d : nd.deps) rb.text += " " + std::to_string(d);
            rb.text += "\n";
        }
    return rb.text;
}

void Executor::free_temps() {
    for (auto &kv : temps) free_device_temps(kv.second);
    temps.clear();
}

} // namespace kunrt
