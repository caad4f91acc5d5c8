/* C ABI of libkunhip.so, the MI355X device runtime of the factor pipeline.
 *
 * The first block has the names, argument meaning and lifetime rules of the
 * reference C API (KunQuant cpp/Kun/CApi.h); the comment of each entry
 * point cites the declaration it replaces. Buffers passed to these entry
 * points are caller-owned host memory, exactly as in the reference: the
 * runtime copies inputs to HBM, runs the stage kernels and copies outputs
 * back before returning.
 *
 * The second block holds extensions that have no counterpart in
 * CApi.h: module introspection (the reference reads its Module struct
 * directly, PyBinding.cpp:182-220), error reporting (the reference throws
 * C++ exceptions), and entry points taking device-resident buffers.
 */
#ifndef KUN_CAPI_H
#define KUN_CAPI_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

#define KUN_API __attribute__((visibility("default")))

typedef void *KunExecutorHandle;
typedef void *KunLibraryHandle;
typedef void *KunModuleHandle;
typedef void *KunBufferNameMapHandle;
typedef void *KunStreamContextHandle;
typedef void *KunStreamPoolHandle;
typedef void *KunBatchGraphHandle;

#define KUN_API_VERSION 1

typedef enum {
    KUN_INIT_NONE = 0,
    KUN_INIT_FILE,
    KUN_INIT_MEMORY,
} KunStateBufferKind;

typedef enum {
    KUN_SUCCESS = 0,
    KUN_INIT_ERROR,
    KUN_INVALID_ARGUMENT,
} KunStatus;

typedef struct {
    size_t version; /* must be KUN_API_VERSION */
    KunStateBufferKind init_kind;
    union {
        const char *path;
        struct {
            const char *buffer;
            size_t size;
        } memory;
    } init;
} KunStreamExtraArgs;

/* ---- reference-compatible entry points ---------------------------------- */

/* CApi.h:60 kunCreateSingleThreadExecutor: an executor owns one HIP stream on
 * the current device and a device buffer pool. */
KUN_API KunExecutorHandle kunCreateSingleThreadExecutor(void);
/* CApi.h:68 kunCreateMultiThreadExecutor: same device executor; numthreads
 * bounds the host threads used for cross-checks and is otherwise ignored. */
KUN_API KunExecutorHandle kunCreateMultiThreadExecutor(int numthreads);
/* CApi.h:76 */
KUN_API void kunDestoryExecutor(KunExecutorHandle ptr);
/* CApi.h:84 kunLoadLibrary: dlopen a device module library built by
 * kunquant_amd.jit.cfake. Returns NULL when loading fails. */
KUN_API KunLibraryHandle kunLoadLibrary(const char *path_or_name);
/* CApi.h:94 kunGetModuleFromLibrary: NULL when the module is not found. */
KUN_API KunModuleHandle kunGetModuleFromLibrary(KunLibraryHandle lib, const char *name);
/* CApi.h:102 */
KUN_API void kunUnloadLibrary(KunLibraryHandle ptr);
/* CApi.h:110-135 buffer name map */
KUN_API KunBufferNameMapHandle kunCreateBufferNameMap(void);
KUN_API void kunDestoryBufferNameMap(KunBufferNameMapHandle ptr);
KUN_API void kunSetBufferNameMap(KunBufferNameMapHandle ptr, const char *name, float *buffer);
KUN_API void kunEraseBufferNameMap(KunBufferNameMapHandle ptr, const char *name);
/* CApi.h:156 kunRunGraph: computes rows [cur_time, cur_time+length) of the
 * inputs (total_time rows each); outputs get `length` rows. Errors are
 * reported through kunGetLastError (the reference lets exceptions escape). */
KUN_API void kunRunGraph(KunExecutorHandle exec, KunModuleHandle m, KunBufferNameMapHandle buffers,
                         size_t num_stocks, size_t total_time, size_t cur_time, size_t length);
/* CApi.h:165 kunCreateStream: NULL on failure. */
KUN_API KunStreamContextHandle kunCreateStream(KunExecutorHandle exec, KunModuleHandle m,
                                               size_t num_stocks);
/* CApi.h:182 kunCreateStreamEx */
KUN_API KunStatus kunCreateStreamEx(KunExecutorHandle exec, KunModuleHandle m, size_t num_stocks,
                                    const KunStreamExtraArgs *extra_args,
                                    KunStreamContextHandle *out_handle);
/* CApi.h:209 kunStreamSerializeStates (argument checks of CApi.cpp:140-164):
 * with KUN_INIT_MEMORY, a NULL `size`, or a NULL buffer with *size != 0, is
 * KUN_INVALID_ARGUMENT; a buffer smaller than the state (a NULL buffer with
 * *size == 0 is the size query) sets *size to the required size and returns
 * KUN_INIT_ERROR. */
KUN_API KunStatus kunStreamSerializeStates(KunStreamContextHandle context,
                                           KunStateBufferKind dump_kind, char *path_or_buffer,
                                           size_t *size);
/* CApi.h:224 */
KUN_API size_t kunQueryBufferHandle(KunStreamContextHandle context, const char *name);
/* CApi.h:235: host pointer to the newest row, valid until the next push/run */
KUN_API const float *kunStreamGetCurrentBuffer(KunStreamContextHandle context, size_t handle);
/* CApi.h:245 */
KUN_API void kunStreamPushData(KunStreamContextHandle context, size_t handle, const float *buffer);
/* CApi.h:253 */
KUN_API void kunStreamRun(KunStreamContextHandle context);
/* CApi.h:259 */
KUN_API void kunDestoryStream(KunStreamContextHandle context);

/* ---- extensions (no counterpart in CApi.h) ------------------------------ */

/* Last error message of the calling thread ("" when none). */
KUN_API const char *kunGetLastError(void);
/* kunRunGraph with a status: 0 on success, nonzero with kunGetLastError set. */
KUN_API int kunRunGraphEx(KunExecutorHandle exec, KunModuleHandle m, KunBufferNameMapHandle buffers,
                          size_t num_stocks, size_t total_time, size_t cur_time, size_t length);
/* Same computation on device-resident buffers: the name map holds device
 * pointers in the module's layouts. Kernels are enqueued on `hip_stream`
 * (NULL: the executor's stream); the call returns without synchronising.
 * Temporaries are kept per (module, shape), so repeated calls with the same
 * shapes allocate nothing after the first; a call waits on its stream for
 * the previous call that used the same temporaries, on whatever stream. */
KUN_API int kunRunGraphDevice(KunExecutorHandle exec, KunModuleHandle m, KunBufferNameMapHandle buffers,
                              size_t num_stocks, size_t total_time, size_t cur_time, size_t length,
                              void *hip_stream);
/* How row-chunked kernels split the rows of a batch run: about
 * `target_workgroups` workgroups per kernel, each chunk at least
 * `halo_factor` times the history it recomputes and `min_rows` rows long.
 * Results do not depend on it. Drops cached device temporaries. */
KUN_API int kunExecutorSetChunking(KunExecutorHandle exec, long target_workgroups, long halo_factor,
                                  long min_rows);
/* Frees the temporaries the executor keeps for device-resident runs (one
 * set per module, shape and job position); waits for the device first.
 * No counterpart in the reference, whose executor allocates per call. */
KUN_API int kunExecutorReleaseTemps(KunExecutorHandle exec);
/* Chunking for one (module, num_stocks, length) only, e.g. as picked by an
 * autotuner; other shapes keep the executor's policy. */
KUN_API int kunExecutorSetChunkingFor(KunExecutorHandle exec, KunModuleHandle m, size_t num_stocks,
                                     size_t length, long target_workgroups, long halo_factor, long min_rows);
/* Independent jobs on device-resident buffers enqueued together (job j:
 * mods[j], bufs[j] holding device pointers, num_stocks[j], ...): their
 * kernels interleave over the executor's lanes. Each job slot keeps its own
 * temporaries. Returns without synchronising. */
KUN_API int kunRunGraphBatchDevice(KunExecutorHandle exec, int njobs, const KunModuleHandle *mods,
                                   const KunBufferNameMapHandle *bufs, const size_t *num_stocks,
                                   const size_t *total_time, const size_t *cur_time, const size_t *length,
                                   void *hip_stream);
/* The same batch as HIP graphs (with temporaries of their own), replayed by
 * kunLaunchGraphBatch: the jobs split into groups of about equal work, as
 * many as the executor has lanes, one graph per group, each launched once
 * per replay on a stream of the batch's own (forked from and joined into
 * the replay stream). The graphs are built node by node from the jobs'
 * launch DAGs (no stream capture, no events inside them). The buffers'
 * device pointers are baked in. NULL on failure (kunGetLastError). */
KUN_API KunBatchGraphHandle kunCaptureGraphBatch(KunExecutorHandle exec, int njobs, const KunModuleHandle *mods,
                                                 const KunBufferNameMapHandle *bufs, const size_t *num_stocks,
                                                 const size_t *total_time, const size_t *cur_time,
                                                 const size_t *length);
/* One replay on `hip_stream` (NULL: the executor's stream). A replay waits
 * for the previous replay of the same graph, whatever its stream (they
 * share the graph's temporaries). Returns without synchronising. */
KUN_API int kunLaunchGraphBatch(KunBatchGraphHandle g, void *hip_stream);
/* Waits for the last replay, then frees the graph and its temporaries. */
KUN_API void kunDestroyGraphBatch(KunBatchGraphHandle g);
/* Host-only description (no device calls) of the graph kunCaptureGraphBatch
 * builds for these jobs, with `lanes` stage lanes in the launch plan: one
 * line per graph node "<id> kernel|empty <job> <stage> <kernel> <grid>
 * <block> : <dependency ids>", then one per launch-DAG node "dag <job>
 * <node> <stage> <kernel> : <dependency nodes>". With a NULL buffer (or
 * one smaller than the text) *size receives the bytes needed. */
KUN_API int kunDescribeGraphBatch(int njobs, const KunModuleHandle *mods, const size_t *num_stocks,
                                  const size_t *total_time, const size_t *cur_time, const size_t *length, int lanes,
                                  char *buffer, size_t *size);
/* Bytes of the fixed device arena of kunRunGraphDeviceNoAlloc (per device). */
#define KUN_NOALLOC_SCRATCH_BYTES (64u << 20)
/* No counterpart in the reference: for callers that cannot allocate device
 * memory (code under HIP graph capture, serving loops, harnesses that forbid
 * hipMalloc). kunRunGraphDevice without any allocation: the run's
 * temporaries, stage-kernel scratch and row sums live in one fixed arena of
 * KUN_NOALLOC_SCRATCH_BYTES per device, held by the library from the first
 * call on a device until the process ends. A run whose memory plan needs
 * more fails with an error (kunGetLastError) and enqueues nothing. Kernels
 * go in launch order on `hip_stream` (NULL: the executor's stream); returns
 * without synchronising. Every call on a device shares that arena: calls
 * that overlap on one device race (order them on one stream). */
KUN_API int kunRunGraphDeviceNoAlloc(KunExecutorHandle exec, KunModuleHandle m, KunBufferNameMapHandle buffers,
                                     size_t num_stocks, size_t total_time, size_t cur_time, size_t length,
                                     void *hip_stream);
/* Load serial of the library the module handle was obtained from: every
 * successful kunLoadLibrary gets a new one. The executor's caches (device
 * temporaries, chunking, stage order, timings) key by (module, serial), so a
 * library loaded where an unloaded one was never meets the old entries. */
KUN_API unsigned long long kunModuleLoadSerial(KunModuleHandle m);
/* Blocks until the executor's stream is idle. */
KUN_API int kunExecutorSynchronize(KunExecutorHandle exec);

/* ---- corrWith / aggregrate (RunGraph.hpp:15-19, 55-75; the reference
 * exposes them to Python only, PyBinding.cpp:396-508) --------------------
 * kunCorrWith: out[i][t - cur_time] = correlation over the stocks of row t of
 * inputs[i] (or of its rank, rank_inputs != 0) with row t of corr_with,
 * float; layout 0 STs ([S/8][T][8]), 1 TS ([T][S]). Host buffers. */
KUN_API int kunCorrWith(KunExecutorHandle exec, int layout, int rank_inputs, int num_inputs,
                        const float *const *inputs, const float *corr_with, float *const *outs, size_t num_stocks,
                        size_t total_time, size_t cur_time, size_t length);
/* device addresses, enqueued on hip_stream (NULL: the executor's stream) */
KUN_API int kunCorrWithDevice(KunExecutorHandle exec, int layout, int rank_inputs, int num_inputs,
                              const float *const *inputs, const float *corr_with, float *const *outs,
                              size_t num_stocks, size_t total_time, size_t cur_time, size_t length, void *hip_stream);
/* kunAggregrate: per aggregation i, rows cur_time .. cur_time + length - 1 of
 * inputs[i] ([T][S], TS) are split into groups where labels[i] ([T])
 * changes; outs[i * 7 + k] (k: sum, min, max, first, last, count, mean; NULL
 * = not wanted) receive [groups][S]. dtype 0 float, 1 double. Host buffers. */
KUN_API int kunAggregrate(KunExecutorHandle exec, int num_aggregrations, const void *const *inputs,
                          const void *const *labels, int dtype, void *const *outs, size_t num_stocks,
                          size_t total_time, size_t cur_time, size_t length);
/* device addresses; every output holds out_rows rows (more groups: error) */
KUN_API int kunAggregrateDevice(KunExecutorHandle exec, int num_aggregrations, const void *const *inputs,
                                const void *const *labels, int dtype, void *const *outs, size_t out_rows,
                                size_t num_stocks, size_t total_time, size_t cur_time, size_t length,
                                void *hip_stream);

/* ---- session pools (no counterpart in CApi.h) ---------------------------
 * Many stream sessions of one STREAM module resident in HBM side by side.
 * A pool has `capacity_stocks` lanes; a session takes a 64-lane aligned
 * range at admission (first fit) and returns it at release, so admission is
 * by state footprint: kunStreamPoolBytesPerStock bytes of rows and state per
 * lane. A cycle ticks a set of idle sessions once each (every session keeps
 * its own tick count), in a fixed number of launches whatever the number of
 * sessions; up to `max_inflight` cycles with disjoint sessions run at once.
 * Per session, results equal those of a stream context of the same module
 * fed the same rows. A pool is not thread-safe. */
KUN_API size_t kunStreamPoolBytesPerStock(KunModuleHandle m);
/* NULL on failure (kunGetLastError) */
KUN_API KunStreamPoolHandle kunCreateStreamPool(KunExecutorHandle exec, KunModuleHandle m, size_t capacity_stocks,
                                                int max_sessions, int max_inflight);
KUN_API void kunDestroyStreamPool(KunStreamPoolHandle pool);
KUN_API size_t kunStreamPoolFreeStocks(KunStreamPoolHandle pool);
/* 0: admitted, *session set; 2: no room now (retry after a release); 1: error */
KUN_API int kunStreamPoolAdmit(KunStreamPoolHandle pool, size_t num_stocks, int *session);
/* waits for the session's cycle in flight, if any */
KUN_API int kunStreamPoolRelease(KunStreamPoolHandle pool, int session);
/* One tick of each of `nsessions` sessions (none in a cycle in flight, none
 * twice). input_rows[j * ninputs + q] is the host row of module input q (in
 * buffer order) of sessions[j]; the rows are copied before the call
 * returns. Enqueues the cycle and returns its id in *cycle. */
KUN_API int kunStreamPoolSubmit(KunStreamPoolHandle pool, int nsessions, const int *sessions,
                                const void *const *input_rows, unsigned long long *cycle);
/* 1: cycle finished, 0: running (wait = 0 only), -1: error */
KUN_API int kunStreamPoolDone(KunStreamPoolHandle pool, unsigned long long cycle, int wait);
/* current row of a buffer of a session: host copy (waits for the session's
 * cycle) or device address (valid until the session's next cycle) */
KUN_API int kunStreamPoolRead(KunStreamPoolHandle pool, int session, size_t buffer_handle, void *host_row);
KUN_API const void *kunStreamPoolRowDevice(KunStreamPoolHandle pool, int session, size_t buffer_handle);
/* ticks submitted for the session so far; -1 on error */
KUN_API long kunStreamPoolTicks(KunStreamPoolHandle pool, int session);
/* Results to host inside each cycle: every later cycle ends by copying the
 * rows of these buffers (handles as kunQueryBufferHandle's) of its sessions
 * to pinned host memory, so a finished cycle's results are in host memory
 * when kunStreamPoolDone reports it. nbuffers = 0 turns it off. Waits for
 * cycles in flight. */
KUN_API int kunStreamPoolSetReadback(KunStreamPoolHandle pool, int nbuffers, const size_t *buffer_handles);
/* 1: the cycle's phases are done (its sessions' rows are final in HBM; the
 * copy to host may still run), 0: running (wait = 0 only), -1: error */
KUN_API int kunStreamPoolComputed(KunStreamPoolHandle pool, unsigned long long cycle, int wait);
/* Host address of a session's row of a readback buffer from cycle `cycle`
 * (waits for the cycle). Valid until max_inflight more cycles have been
 * submitted after it (slots are reused oldest first); NULL on error. */
KUN_API const void *kunStreamPoolResult(KunStreamPoolHandle pool, unsigned long long cycle, int session,
                                        size_t buffer_handle);
/* Per-phase timing of later cycles (enable = 1): kunStreamPoolPhaseTimes
 * fills ms[0] = input staging (copy + scatter), ms[1 .. nphase] = the
 * module's phases in launch order, ms[nphase + 1] = readback (0 without),
 * waiting for the cycle; returns nphase + 2, or -1 on error. */
KUN_API int kunStreamPoolProfile(KunStreamPoolHandle pool, int enable);
KUN_API int kunStreamPoolPhaseTimes(KunStreamPoolHandle pool, unsigned long long cycle, float *ms, int n);
/* A session resumed from a serialised stream state in the reference's byte
 * layout (kunStreamSerializeStates of either library, same module and
 * stocks). Return codes as kunStreamPoolAdmit; a bad blob is error 1 with
 * the reference's messages. */
KUN_API int kunStreamPoolAdmitState(KunStreamPoolHandle pool, size_t num_stocks, const char *state, size_t state_len,
                                    int *session);
/* The session's state in the reference's byte layout (waits for its cycle).
 * *size = 0: *size receives the length only; otherwise buffer holds *size
 * bytes and receives the blob, *size its length (error 1 if too small). */
KUN_API int kunStreamPoolSerializeStates(KunStreamPoolHandle pool, int session, char *buffer, size_t *size);

/* Module introspection. dtype: 0 float, 1 double. layouts: 0 STs, 1 TS, 2 STREAM. */
KUN_API int kunModuleDtype(KunModuleHandle m);
KUN_API int kunModuleInputLayout(KunModuleHandle m);
KUN_API int kunModuleOutputLayout(KunModuleHandle m);
KUN_API int kunModuleBlockingLen(KunModuleHandle m);
KUN_API int kunModuleAligned(KunModuleHandle m);
KUN_API int kunModuleNumBuffers(KunModuleHandle m);
/* kind: 0 input, 1 output, 2 temporary */
KUN_API const char *kunModuleBufferName(KunModuleHandle m, int idx);
KUN_API int kunModuleBufferKind(KunModuleHandle m, int idx);
KUN_API int kunModuleBufferUnreliable(KunModuleHandle m, int idx);
KUN_API int kunModuleBufferWindow(KunModuleHandle m, int idx);
KUN_API int kunModuleNumStages(KunModuleHandle m);
/* KUNPROG text describing the module's stage IR */
KUN_API const char *kunModuleProgram(KunModuleHandle m);
/* stage table: kind 0 time-series, 1 Rank, 2 Scale; inputs/outputs are buffer indices */
KUN_API const char *kunModuleStageName(KunModuleHandle m, int stage);
KUN_API int kunModuleStageKind(KunModuleHandle m, int stage);
KUN_API int kunModuleStageNumInputs(KunModuleHandle m, int stage);
KUN_API int kunModuleStageInput(KunModuleHandle m, int stage, int j);
KUN_API int kunModuleStageNumOutputs(KunModuleHandle m, int stage);
KUN_API int kunModuleStageOutput(KunModuleHandle m, int stage, int j);
/* kernels a stage launches per batch run (cross-sectional stages: 1), and
 * the plan of one kernel of a time-series stage: kind 0 row-chunked, 1
 * sequential over rows; ncomp independent op groups; halo rows recomputed
 * before each chunk; static LDS bytes per workgroup; distinct HBM values
 * the kernel reads or writes per (stock, row). */
KUN_API int kunModuleStageNumKernels(KunModuleHandle m, int stage);
KUN_API int kunModuleKernelInfo(KunModuleHandle m, int stage, int kern, int *kind, int *ncomp, int *halo,
                                int *lds_bytes, int *io_values);
/* Device memory a batch run of `length` rows over `num_stocks` stocks holds
 * besides its inputs and outputs (temporaries, values handed between a
 * stage's kernels, HBM ring columns), under the default row chunking; the
 * [rows][stocks] slots it reuses are counted in *num_slots (may be null).
 * Returns 0 on error. Host-only: touches no GPU. */
KUN_API size_t kunModuleDeviceBytes(KunModuleHandle m, size_t num_stocks, size_t length, int *num_slots);
/* Kernel timing with HIP events recorded on the stream each kernel is
 * launched on. While enabled, every batch run records an event pair around
 * each kernel; kunExecutorStageTime synchronises and returns the accumulated
 * milliseconds of all kernels of one stage and the number of runs;
 * kunExecutorKernelTime the same for one kernel of the stage. */
KUN_API int kunExecutorEnableProfiling(KunExecutorHandle exec, int on);
KUN_API int kunExecutorStageTime(KunExecutorHandle exec, KunModuleHandle m, int stage, double *total_ms,
                                 long *count);
KUN_API int kunExecutorKernelTime(KunExecutorHandle exec, KunModuleHandle m, int stage, int kern, double *total_ms,
                                  long *count);
/* device pointer of the newest row of a stream buffer (num_stocks elements) */
KUN_API const void *kunStreamGetCurrentDeviceBuffer(KunStreamContextHandle context, size_t handle);
/* number of rows consumed by the stream so far */
KUN_API size_t kunStreamTicks(KunStreamContextHandle context);

#ifdef __cplusplus
}
#endif

#endif
