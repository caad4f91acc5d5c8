// Internal interfaces of the device runtime (runtime.hip, stream.hip, capi.cpp).
#pragma once
#include <hip/hip_runtime.h>

#include <map>
#include <memory>
#include <mutex>
#include <stdexcept>
#include <string>
#include <tuple>
#include <unordered_map>
#include <vector>

#include "kun_module_abi.h"
#include "lane_sync.hpp"

#ifndef KUN_MAX_COMPS
#define KUN_MAX_COMPS 64
#endif

namespace kunrt {

// Stream state in the reference's byte layout (stream.hip). WireHost: host
// copies of one stream's rows ([S] per module buffer) and stage states
// ([slot][S] T columns, then int32 columns); WireView: the stream's module,
// stocks, tick count, tick count at its last resume and the per-wire-buffer
// positions recorded then.
struct WireHost {
    std::vector<std::vector<char>> rows, state;
};
struct WireView {
    const KunModuleDesc *m;
    size_t S;
    long tick, tick0;
    std::vector<uint64_t> *pos0;
};
void wire_serialize(const WireView &v, WireHost &h, std::string &out);
// throws "Failed to read initial stream buffer" / "Failed to deserialize state buffer"
void wire_deserialize(const WireView &v, WireHost &h, const char *p, const char *end);

struct Error : std::runtime_error {
    using std::runtime_error::runtime_error;
};

#define HIPCHECK(x)                                                                              \
    do {                                                                                         \
        hipError_t e_ = (x);                                                                     \
        if (e_ != hipSuccess) throw ::kunrt::Error(std::string(#x) + ": " + hipGetErrorString(e_)); \
    } while (0)

void set_error(const std::string &msg);
const char *last_error();

using NameMap = std::unordered_map<std::string, void *>;

// exact-size caching device allocator
struct Pool {
    std::mutex mu;
    std::multimap<size_t, void *> free_;
    void *get(size_t bytes);
    void put(void *p, size_t bytes);
    ~Pool();
};

// Device memory of a batch run, one arena: per kernel, its row chunking and
// HBM columns; then the [rows][stocks] slots that hold the module's
// temporaries and the values handed between a stage's kernels (a slot is
// reused once every kernel touching its previous value precedes the new
// writer in the launch DAG).
struct BatchLayout {
    struct Kern {
        size_t scratch = 0;       // arena offset of the kernel's scratch
        unsigned long blocks = 0; // workgroups of the flat grid
        int nz = 0;               // launch table (kun::StageArgs z* fields)
        int zcomp[KUN_MAX_COMPS] = {}, zfirst[KUN_MAX_COMPS + 1] = {}, zchunk[KUN_MAX_COMPS] = {},
            znch[KUN_MAX_COMPS] = {};
        long zscr[KUN_MAX_COMPS] = {};
    };
    std::vector<std::vector<Kern>> kern;
    std::vector<size_t> cs_scratch;          // per Scale stage: arena offset of its row sums
    std::vector<long> buf_off;                // per module buffer: arena offset of a temporary, else -1
    std::vector<std::vector<size_t>> mat_off; // per stage, per materialised value: arena offset
    size_t slot_bytes = 0;
    int nslots = 0;
    size_t bytes = 0;
    // launch DAG: one node per stage kernel, in a topological order
    struct Node {
        int stage = 0, kern = 0, lane = 0;
        std::vector<int> deps;
        double est = 0; // estimated microseconds
    };
    std::vector<Node> nodes;
    // cross-lane synchronisation of one run of these nodes on `sync_lanes`
    // lanes (plan_lane_sync), computed at the first run
    LaneSync sync;
    int sync_lanes = -1;
};
// How PAR kernels split their rows (defaults overridable by the environment
// variables KUN_TARGET_WORKGROUPS, KUN_HALO_FACTOR, KUN_MIN_CHUNK_ROWS).
struct ChunkPolicy {
    long target_workgroups = 32768;
    long halo_factor = 3;
    long min_rows = 64;
    // a kernel left with fewer workgroups than this (16 per CU) splits its
    // components' rows further, down to chunks of one halo (at least 16
    // rows): the large-halo kernels otherwise run a few hundred long waves
    long min_workgroups = 4096;
    static ChunkPolicy from_env();
};
BatchLayout batch_layout(const KunModuleDesc *m, size_t S, size_t L, const ChunkPolicy &pol, int lanes);

struct DeviceTemps {
    std::vector<void *> bufs; // per module buffer: the temporary's address in the arena, else null
    char *arena = nullptr;
    BatchLayout lay;
    // recorded on the caller's stream once a run's lanes have joined; the
    // next run using this arena waits for it first, whatever its stream
    hipEvent_t done = nullptr;
    bool armed = false;
};

// An executor: one device, one HIP stream that calls are ordered on, a few
// lanes (streams) that independent stages of a batch run spread over, a
// buffer pool.
struct Executor {
    int device = 0;
    hipStream_t stream = nullptr;
    std::vector<hipStream_t> lanes;
    std::vector<hipEvent_t> lane_done, stage_done;
    hipEvent_t fork = nullptr;
    void ensure_stage_events(int n);
    Pool pool;
    std::mutex mu;
    ChunkPolicy chunking;
    // temporaries of device-resident runs, per (module, load serial, stocks,
    // rows, job slot); slot 0: single runs, j + 1: job j of a batch
    std::map<std::tuple<const void *, uint64_t, size_t, size_t, int>, DeviceTemps> temps;
    // chunking chosen for one (module, load serial, stocks, rows), e.g. by the autotuner
    std::map<std::tuple<const void *, uint64_t, size_t, size_t>, ChunkPolicy> tuned;
    ChunkPolicy policy_for(const KunModuleDesc *m, size_t S, size_t L) const;
    // per-stage kernel timing (bench / diagnostics)
    bool profiling = false;
    struct Pending {
        const void *module;
        uint64_t serial;
        int stage, kern;
        hipEvent_t start, stop;
    };
    std::vector<Pending> pending;
    std::vector<hipEvent_t> event_pool;
    // (module, load serial, stage, kernel of the stage) -> (milliseconds, launches)
    std::map<std::tuple<const void *, uint64_t, int, int>, std::pair<double, long>> stage_ms;
    hipEvent_t take_event();
    void collect_profile();
    Executor();
    ~Executor();
    void free_temps();
};

struct Library {
    void *handle = nullptr;
    std::string path;
    uint64_t serial = 0; // load serial: distinct for every successful load in the process
    static std::shared_ptr<Library> load(const char *path);
    const KunModuleDesc *get_module(const char *name);
    ~Library();
};
// Load serial of the library a module descriptor was last obtained from
// (0 for descriptors that did not come through Library::get_module). A
// library loaded after another was unloaded can place its descriptors at
// the old addresses; every cache keyed by a descriptor also keys by this.
uint64_t module_serial(const KunModuleDesc *m);

std::string module_symbol(const std::string &name);
const std::vector<int> &stage_order(const KunModuleDesc *m);
size_t elem_size(const KunModuleDesc *m);
size_t buffer_elems(const KunModuleDesc *m, int b, size_t S, size_t T, size_t L);
void run_graph_host(Executor &ex, const KunModuleDesc *m, const NameMap &bufs, size_t S, size_t T,
                    size_t start, size_t L);
// argument checks of a batch run (throws)
void check_batch_module(const KunModuleDesc *m, size_t S, size_t T, size_t start, size_t L);
// every node of `lay` in launch order on one stream; dev: device pointers of
// the module's buffers, arena: the memory `lay` describes; prof: per-kernel
// event timing (or null)
void enqueue_batch_serial(const KunModuleDesc *m, const std::vector<void *> &dev, const BatchLayout &lay,
                          char *arena, size_t S, size_t T, size_t start, size_t L, hipStream_t st,
                          Executor *prof = nullptr);
void run_graph_device(Executor &ex, const KunModuleDesc *m, const NameMap &bufs, size_t S, size_t T,
                      size_t start, size_t L, hipStream_t st);

// one job of a batched device run
struct BatchJob {
    const KunModuleDesc *m;
    const NameMap *bufs;
    size_t S, T, start, L;
};
void run_jobs_device(Executor &ex, const std::vector<BatchJob> &jobs, hipStream_t st);

// a batch of jobs as HIP graphs with temporaries of their own: the jobs
// split into groups of about equal work, one graph per group, replayed side
// by side on the batch's own streams (forked from and joined into the
// caller's stream)
struct CapturedBatch {
    int device = 0;
    std::shared_ptr<Executor> ex; // its stream is the default replay stream
    std::vector<hipGraph_t> graphs;
    std::vector<hipGraphExec_t> execs;
    std::vector<hipStream_t> streams; // one per graph
    std::vector<hipEvent_t> joined;   // one per graph, recorded after its replay
    hipEvent_t fork = nullptr;
    std::vector<DeviceTemps> temps;
    // recorded after each replay; a replay on another stream waits for it
    hipEvent_t done = nullptr;
    bool armed = false;
    ~CapturedBatch();
};
// Builds a graph node by node. Node handles are opaque; deps are handles
// returned earlier by the same builder.
struct GraphBuilder {
    virtual hipError_t add_kernel(const void *func, dim3 grid, dim3 block, void **args,
                                  const std::vector<void *> &deps, void **node) = 0;
    virtual hipError_t add_empty(const std::vector<void *> &deps, void **node) = 0;
    virtual void tag(int job, int stage, int kern) { (void)job, (void)stage, (void)kern; }
    virtual ~GraphBuilder() = default;
};
void build_jobs_graph(GraphBuilder &gb, const std::vector<BatchJob> &jobs,
                      const std::vector<const DeviceTemps *> &temps);
CapturedBatch *capture_jobs(std::shared_ptr<Executor> ex, const std::vector<BatchJob> &jobs);
void launch_captured(CapturedBatch &cb, hipStream_t st);
// text description of the graph capture_jobs builds for these jobs (host
// only): nodes with their dependencies, and the jobs' launch DAGs
std::string describe_jobs_graph(const std::vector<BatchJob> &jobs, int lanes);
DeviceTemps make_temps(Executor &ex, const KunModuleDesc *m, size_t S, size_t T, size_t L);
void free_device_temps(DeviceTemps &t);

// ---- stream contexts ----------------------------------------------------
struct StreamCtx;
StreamCtx *stream_create(std::shared_ptr<Executor> ex, const KunModuleDesc *m, size_t S,
                         const char *state, size_t state_len, bool from_file);
void stream_destroy(StreamCtx *c);
size_t stream_query(StreamCtx *c, const char *name);
void stream_push(StreamCtx *c, size_t handle, const void *host_row);
void stream_run(StreamCtx *c);
const void *stream_current(StreamCtx *c, size_t handle);
const void *stream_current_device(StreamCtx *c, size_t handle);
size_t stream_ticks(StreamCtx *c);
std::string stream_serialize(StreamCtx *c);

// ---- corrWith / aggregrate (row_reduce.hip) ------------------------------
// device: the pointers are device addresses on `st` (no synchronisation);
// otherwise host buffers (copied, synchronous)
void corr_with(Executor &ex, bool device, int sts, bool rank_inputs, int n, const float *const *inputs,
               const float *target, float *const *outs, size_t S, size_t T, size_t start, size_t L, int blk,
               hipStream_t st);
// outs: [n][7] (sum, min, max, first, last, count, mean; null = not wanted);
// device outputs hold out_rows rows each
void aggregate(Executor &ex, bool device, int n, const void *const *inputs, const void *const *labels, int dtype,
               void *const *outs, size_t S, size_t T, size_t start, size_t L, size_t out_rows, hipStream_t st);

// ---- session pools (stream_pool.hip) ------------------------------------
struct StreamPool;
size_t pool_bytes_per_stock(const KunModuleDesc *m);
StreamPool *pool_create(std::shared_ptr<Executor> ex, const KunModuleDesc *m, size_t capacity, int max_sessions,
                        int max_inflight);
void pool_destroy(StreamPool *p);
size_t pool_free_stocks(StreamPool *p);
int pool_admit(StreamPool *p, size_t S);
void pool_release(StreamPool *p, int sid);
unsigned long long pool_submit(StreamPool *p, int n, const int *sessions, const void *const *inputs);
int pool_done(StreamPool *p, unsigned long long cycle, bool wait);
const void *pool_row_device(StreamPool *p, int sid, size_t b);
void pool_read(StreamPool *p, int sid, size_t b, void *host);
long pool_ticks(StreamPool *p, int sid);
void pool_set_readback(StreamPool *p, int n, const size_t *bufs);
const void *pool_result(StreamPool *p, unsigned long long cycle, int sid, size_t b);
void pool_profile(StreamPool *p, bool on);
int pool_computed(StreamPool *p, unsigned long long cycle, bool wait);
int pool_phase_times(StreamPool *p, unsigned long long cycle, float *ms, int n);
int pool_admit_state(StreamPool *p, size_t S, const char *state, size_t len);
std::string pool_serialize(StreamPool *p, int sid);

} // namespace kunrt
