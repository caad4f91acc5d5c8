// extern "C" entry points of libkunhip.so (declared in include/kun_capi.h).
// Handles follow the reference's conventions (cpp/Kun/CApi.cpp:11-189):
// executors and libraries are heap-boxed shared_ptrs, module handles are
// borrowed descriptor pointers that stay valid while the library is loaded.
#include <stdint.h>
#include <string.h>

#include <fstream>
#include <iterator>
#include <memory>
#include <string>
#include <unordered_map>

#include "../../include/kun_capi.h"
#include "kun_rt.hpp"

using namespace kunrt;

namespace {
std::shared_ptr<Executor> *unwrap_exec(KunExecutorHandle h) { return (std::shared_ptr<Executor> *)h; }
std::shared_ptr<Library> *unwrap_lib(KunLibraryHandle h) { return (std::shared_ptr<Library> *)h; }
NameMap *unwrap_map(KunBufferNameMapHandle h) { return (NameMap *)h; }
const KunModuleDesc *unwrap_mod(KunModuleHandle h) { return (const KunModuleDesc *)h; }

template <typename F> int guarded(F &&f) {
    try {
        set_error("");
        f();
        return 0;
    } catch (const std::exception &e) {
        set_error(e.what());
        return 1;
    }
}
} // namespace

extern "C" {

KUN_API const char *kunGetLastError(void) { return last_error(); }

KUN_API KunExecutorHandle kunCreateSingleThreadExecutor(void) {
    std::shared_ptr<Executor> *h = nullptr;
    guarded([&] { h = new std::shared_ptr<Executor>(std::make_shared<Executor>()); });
    return h;
}

KUN_API KunExecutorHandle kunCreateMultiThreadExecutor(int numthreads) {
    (void)numthreads;
    return kunCreateSingleThreadExecutor();
}

KUN_API void kunDestoryExecutor(KunExecutorHandle ptr) { delete unwrap_exec(ptr); }

KUN_API int kunExecutorSetChunking(KunExecutorHandle exec, long target_workgroups, long halo_factor,
                                  long min_rows) {
    return guarded([&] {
        if (target_workgroups < 1 || halo_factor < 0 || min_rows < 1) throw Error("bad chunking parameters");
        Executor &ex = **unwrap_exec(exec);
        std::lock_guard<std::mutex> g(ex.mu);
        HIPCHECK(hipStreamSynchronize(ex.stream));
        ex.free_temps();
        ex.chunking.target_workgroups = target_workgroups;
        ex.chunking.halo_factor = halo_factor;
        ex.chunking.min_rows = min_rows;
    });
}

KUN_API int kunExecutorReleaseTemps(KunExecutorHandle exec) {
    return guarded([&] {
        Executor &ex = **unwrap_exec(exec);
        std::lock_guard<std::mutex> g(ex.mu);
        HIPCHECK(hipSetDevice(ex.device));
        HIPCHECK(hipDeviceSynchronize());
        ex.free_temps();
    });
}

KUN_API int kunExecutorSetChunkingFor(KunExecutorHandle exec, KunModuleHandle m, size_t num_stocks,
                                     size_t length, long target_workgroups, long halo_factor, long min_rows) {
    return guarded([&] {
        if (target_workgroups < 1 || halo_factor < 0 || min_rows < 1) throw Error("bad chunking parameters");
        Executor &ex = **unwrap_exec(exec);
        std::lock_guard<std::mutex> g(ex.mu);
        HIPCHECK(hipStreamSynchronize(ex.stream));
        ChunkPolicy p;
        p.target_workgroups = target_workgroups;
        p.halo_factor = halo_factor;
        p.min_rows = min_rows;
        const uint64_t serial = module_serial(unwrap_mod(m));
        ex.tuned[std::make_tuple((const void *)m, serial, num_stocks, length)] = p;
        // drop cached temporaries of that module and shape: their layout changes
        for (auto it = ex.temps.begin(); it != ex.temps.end();) {
            if (std::get<0>(it->first) == (const void *)m && std::get<1>(it->first) == serial &&
                std::get<2>(it->first) == num_stocks && std::get<3>(it->first) == length) {
                free_device_temps(it->second);
                it = ex.temps.erase(it);
            } else {
                ++it;
            }
        }
    });
}

static std::vector<BatchJob> make_jobs(int njobs, const KunModuleHandle *mods, const KunBufferNameMapHandle *bufs,
                                       const size_t *num_stocks, const size_t *total_time, const size_t *cur_time,
                                       const size_t *length) {
    if (njobs < 1) throw Error("a batch needs at least one job");
    std::vector<BatchJob> jobs(njobs);
    for (int j = 0; j < njobs; j++)
        jobs[j] = BatchJob{unwrap_mod(mods[j]), unwrap_map(bufs[j]), num_stocks[j], total_time[j], cur_time[j],
                           length[j]};
    return jobs;
}

KUN_API int kunRunGraphBatchDevice(KunExecutorHandle exec, int njobs, const KunModuleHandle *mods,
                                   const KunBufferNameMapHandle *bufs, const size_t *num_stocks,
                                   const size_t *total_time, const size_t *cur_time, const size_t *length,
                                   void *hip_stream) {
    return guarded([&] {
        run_jobs_device(**unwrap_exec(exec), make_jobs(njobs, mods, bufs, num_stocks, total_time, cur_time, length),
                        (hipStream_t)hip_stream);
    });
}

KUN_API KunBatchGraphHandle kunCaptureGraphBatch(KunExecutorHandle exec, int njobs, const KunModuleHandle *mods,
                                                 const KunBufferNameMapHandle *bufs, const size_t *num_stocks,
                                                 const size_t *total_time, const size_t *cur_time,
                                                 const size_t *length) {
    CapturedBatch *cb = nullptr;
    guarded([&] {
        cb = capture_jobs(*unwrap_exec(exec), make_jobs(njobs, mods, bufs, num_stocks, total_time, cur_time, length));
    });
    return cb;
}

KUN_API int kunLaunchGraphBatch(KunBatchGraphHandle g, void *hip_stream) {
    return guarded([&] {
        CapturedBatch *cb = (CapturedBatch *)g;
        if (!cb) throw Error("null batch graph");
        launch_captured(*cb, (hipStream_t)hip_stream);
    });
}

KUN_API int kunDescribeGraphBatch(int njobs, const KunModuleHandle *mods, const size_t *num_stocks,
                                  const size_t *total_time, const size_t *cur_time, const size_t *length, int lanes,
                                  char *buffer, size_t *size) {
    return guarded([&] {
        if (!size) throw Error("size is required");
        if (njobs < 1) throw Error("a batch needs at least one job");
        // every non-temporary buffer bound to a placeholder address
        std::vector<NameMap> maps(njobs);
        std::vector<BatchJob> jobs(njobs);
        for (int j = 0; j < njobs; j++) {
            const KunModuleDesc *m = unwrap_mod(mods[j]);
            if (!m) throw Error("null module");
            for (int b = 0; b < m->nbuf; b++)
                if (m->bufs[b].kind != KUN_BUF_TEMP) maps[j][m->bufs[b].name] = (void *)(uintptr_t)(4096 * (b + 1));
            jobs[j] = BatchJob{m, &maps[j], num_stocks[j], total_time[j], cur_time[j], length[j]};
        }
        const std::string text = describe_jobs_graph(jobs, lanes);
        if (!buffer || *size < text.size() + 1) {
            *size = text.size() + 1;
            if (buffer) throw Error("buffer too small");
            return;
        }
        memcpy(buffer, text.c_str(), text.size() + 1);
        *size = text.size() + 1;
    });
}

KUN_API void kunDestroyGraphBatch(KunBatchGraphHandle g) { delete (CapturedBatch *)g; }

KUN_API int kunExecutorSynchronize(KunExecutorHandle exec) {
    return guarded([&] { HIPCHECK(hipStreamSynchronize((*unwrap_exec(exec))->stream)); });
}

KUN_API KunLibraryHandle kunLoadLibrary(const char *path_or_name) {
    auto lib = Library::load(path_or_name);
    if (!lib) {
        set_error(std::string("Cannot load library ") + path_or_name);
        return nullptr;
    }
    return new std::shared_ptr<Library>(lib);
}

KUN_API KunModuleHandle kunGetModuleFromLibrary(KunLibraryHandle lib, const char *name) {
    return (KunModuleHandle)(*unwrap_lib(lib))->get_module(name);
}

KUN_API void kunUnloadLibrary(KunLibraryHandle ptr) { delete unwrap_lib(ptr); }

KUN_API unsigned long long kunModuleLoadSerial(KunModuleHandle m) { return module_serial(unwrap_mod(m)); }

KUN_API KunBufferNameMapHandle kunCreateBufferNameMap(void) { return new NameMap(); }
KUN_API void kunDestoryBufferNameMap(KunBufferNameMapHandle ptr) { delete unwrap_map(ptr); }
KUN_API void kunSetBufferNameMap(KunBufferNameMapHandle ptr, const char *name, float *buffer) {
    (*unwrap_map(ptr))[name] = buffer;
}
KUN_API void kunEraseBufferNameMap(KunBufferNameMapHandle ptr, const char *name) {
    unwrap_map(ptr)->erase(name);
}

KUN_API int kunRunGraphEx(KunExecutorHandle exec, KunModuleHandle m, KunBufferNameMapHandle buffers,
                          size_t num_stocks, size_t total_time, size_t cur_time, size_t length) {
    return guarded([&] {
        run_graph_host(**unwrap_exec(exec), unwrap_mod(m), *unwrap_map(buffers), num_stocks, total_time,
                       cur_time, length);
    });
}

KUN_API void kunRunGraph(KunExecutorHandle exec, KunModuleHandle m, KunBufferNameMapHandle buffers,
                         size_t num_stocks, size_t total_time, size_t cur_time, size_t length) {
    if (kunRunGraphEx(exec, m, buffers, num_stocks, total_time, cur_time, length))
        fprintf(stderr, "kunRunGraph: %s\n", last_error());
}

KUN_API int kunRunGraphDevice(KunExecutorHandle exec, KunModuleHandle m, KunBufferNameMapHandle buffers,
                              size_t num_stocks, size_t total_time, size_t cur_time, size_t length,
                              void *hip_stream) {
    return guarded([&] {
        run_graph_device(**unwrap_exec(exec), unwrap_mod(m), *unwrap_map(buffers), num_stocks, total_time,
                         cur_time, length, (hipStream_t)hip_stream);
    });
}

// ---- module introspection ----------------------------------------------
KUN_API int kunModuleDtype(KunModuleHandle m) { return unwrap_mod(m)->dtype; }
KUN_API int kunModuleInputLayout(KunModuleHandle m) { return unwrap_mod(m)->in_layout; }
KUN_API int kunModuleOutputLayout(KunModuleHandle m) { return unwrap_mod(m)->out_layout; }
KUN_API int kunModuleBlockingLen(KunModuleHandle m) { return unwrap_mod(m)->blocking; }
KUN_API int kunModuleAligned(KunModuleHandle m) { return unwrap_mod(m)->aligned; }
KUN_API int kunModuleNumBuffers(KunModuleHandle m) { return unwrap_mod(m)->nbuf; }
KUN_API const char *kunModuleBufferName(KunModuleHandle m, int idx) { return unwrap_mod(m)->bufs[idx].name; }
KUN_API int kunModuleBufferKind(KunModuleHandle m, int idx) { return unwrap_mod(m)->bufs[idx].kind; }
KUN_API int kunModuleBufferUnreliable(KunModuleHandle m, int idx) { return unwrap_mod(m)->bufs[idx].unreliable; }
KUN_API int kunModuleBufferWindow(KunModuleHandle m, int idx) { return unwrap_mod(m)->bufs[idx].window; }
KUN_API int kunModuleNumStages(KunModuleHandle m) { return unwrap_mod(m)->nstage; }
KUN_API const char *kunModuleProgram(KunModuleHandle m) { return unwrap_mod(m)->program; }

KUN_API const char *kunModuleStageName(KunModuleHandle m, int i) { return unwrap_mod(m)->stages[i].name; }
KUN_API int kunModuleStageKind(KunModuleHandle m, int i) { return unwrap_mod(m)->stages[i].kind; }
KUN_API int kunModuleStageNumInputs(KunModuleHandle m, int i) { return unwrap_mod(m)->stages[i].nin; }
KUN_API int kunModuleStageInput(KunModuleHandle m, int i, int j) { return unwrap_mod(m)->stages[i].in[j]; }
KUN_API int kunModuleStageNumOutputs(KunModuleHandle m, int i) { return unwrap_mod(m)->stages[i].nout; }
KUN_API int kunModuleStageOutput(KunModuleHandle m, int i, int j) { return unwrap_mod(m)->stages[i].out[j]; }
KUN_API int kunModuleStageNumKernels(KunModuleHandle m, int i) {
    const KunStageDesc &sd = unwrap_mod(m)->stages[i];
    return sd.kind == KUN_STAGE_TS ? sd.nkern : 1;
}
KUN_API int kunModuleKernelInfo(KunModuleHandle m, int stage, int kern, int *kind, int *ncomp, int *halo,
                                int *lds_bytes, int *io_values) {
    const KunStageDesc &sd = unwrap_mod(m)->stages[stage];
    if (sd.kind != KUN_STAGE_TS || kern < 0 || kern >= sd.nkern) {
        set_error("no such stage kernel");
        return 1;
    }
    const KunKernelDesc &kd = sd.kerns[kern];
    *kind = kd.kind;
    *ncomp = kd.ncomp;
    *halo = kd.halo;
    *lds_bytes = kd.lds_bytes;
    *io_values = kd.io_values;
    return 0;
}

KUN_API size_t kunModuleDeviceBytes(KunModuleHandle m, size_t num_stocks, size_t length, int *num_slots) {
    size_t bytes = 0;
    int rc = guarded([&] {
        BatchLayout lay = batch_layout(unwrap_mod(m), num_stocks, length, ChunkPolicy::from_env(), 4);
        bytes = lay.bytes;
        if (num_slots) *num_slots = lay.nslots;
    });
    return rc ? 0 : bytes;
}

KUN_API int kunExecutorEnableProfiling(KunExecutorHandle exec, int on) {
    return guarded([&] {
        Executor &ex = **unwrap_exec(exec);
        std::lock_guard<std::mutex> g(ex.mu);
        ex.collect_profile();
        ex.profiling = on != 0;
        if (!on) ex.stage_ms.clear();
    });
}

KUN_API int kunExecutorStageTime(KunExecutorHandle exec, KunModuleHandle m, int stage, double *total_ms,
                                 long *count) {
    return guarded([&] {
        Executor &ex = **unwrap_exec(exec);
        std::lock_guard<std::mutex> g(ex.mu);
        ex.collect_profile();
        *total_ms = 0.0;
        *count = 0;
        const uint64_t serial = module_serial(unwrap_mod(m));
        for (auto &kv : ex.stage_ms) {
            if (std::get<0>(kv.first) != (const void *)m || std::get<1>(kv.first) != serial ||
                std::get<2>(kv.first) != stage)
                continue;
            *total_ms += kv.second.first;
            if (std::get<3>(kv.first) == 0) *count = kv.second.second;
        }
    });
}

KUN_API int kunExecutorKernelTime(KunExecutorHandle exec, KunModuleHandle m, int stage, int kern, double *total_ms,
                                  long *count) {
    return guarded([&] {
        Executor &ex = **unwrap_exec(exec);
        std::lock_guard<std::mutex> g(ex.mu);
        ex.collect_profile();
        auto it = ex.stage_ms.find(std::make_tuple((const void *)m, module_serial(unwrap_mod(m)), stage, kern));
        *total_ms = it == ex.stage_ms.end() ? 0.0 : it->second.first;
        *count = it == ex.stage_ms.end() ? 0 : it->second.second;
    });
}

// ---- stream contexts ----------------------------------------------------
KUN_API KunStatus kunCreateStreamEx(KunExecutorHandle exec, KunModuleHandle m, size_t num_stocks,
                                    const KunStreamExtraArgs *extra_args, KunStreamContextHandle *out_handle) {
    if (!out_handle) return KUN_INVALID_ARGUMENT;
    *out_handle = nullptr;
    const char *state = nullptr;
    size_t state_len = 0;
    bool from_file = false;
    if (extra_args) {
        if (extra_args->version != KUN_API_VERSION) return KUN_INVALID_ARGUMENT;
        if (extra_args->init_kind == KUN_INIT_FILE) {
            if (!extra_args->init.path) return KUN_INVALID_ARGUMENT;
            state = extra_args->init.path;
            from_file = true;
        } else if (extra_args->init_kind == KUN_INIT_MEMORY) {
            if (!extra_args->init.memory.buffer) return KUN_INVALID_ARGUMENT;
            state = extra_args->init.memory.buffer;
            state_len = extra_args->init.memory.size;
        } else if (extra_args->init_kind != KUN_INIT_NONE) {
            return KUN_INVALID_ARGUMENT;
        }
    }
    StreamCtx *ctx = nullptr;
    int rc = guarded([&] { ctx = stream_create(*unwrap_exec(exec), unwrap_mod(m), num_stocks, state, state_len, from_file); });
    if (rc) return KUN_INIT_ERROR;
    *out_handle = ctx;
    return KUN_SUCCESS;
}

KUN_API KunStreamContextHandle kunCreateStream(KunExecutorHandle exec, KunModuleHandle m, size_t num_stocks) {
    KunStreamContextHandle h = nullptr;
    if (kunCreateStreamEx(exec, m, num_stocks, nullptr, &h) != KUN_SUCCESS) return nullptr;
    return h;
}

KUN_API KunStatus kunStreamSerializeStates(KunStreamContextHandle context, KunStateBufferKind dump_kind,
                                           char *path_or_buffer, size_t *size) {
    if (dump_kind != KUN_INIT_FILE && dump_kind != KUN_INIT_MEMORY) return KUN_INVALID_ARGUMENT;
    // a memory dump needs `size`, and a buffer whenever *size is not zero
    // (a zero size with no buffer is the size query)
    if (dump_kind == KUN_INIT_MEMORY && (!size || (*size != 0 && !path_or_buffer))) return KUN_INVALID_ARGUMENT;
    std::string blob;
    if (guarded([&] { blob = stream_serialize((StreamCtx *)context); })) return KUN_INIT_ERROR;
    if (dump_kind == KUN_INIT_FILE) {
        if (!path_or_buffer) return KUN_INVALID_ARGUMENT;
        std::ofstream f(path_or_buffer, std::ios::binary);
        if (!f) return KUN_INIT_ERROR;
        f.write(blob.data(), (std::streamsize)blob.size());
        return f ? KUN_SUCCESS : KUN_INIT_ERROR;
    }
    if (!path_or_buffer || *size < blob.size()) {
        *size = blob.size();
        return KUN_INIT_ERROR;
    }
    memcpy(path_or_buffer, blob.data(), blob.size());
    *size = blob.size();
    return KUN_SUCCESS;
}

KUN_API size_t kunQueryBufferHandle(KunStreamContextHandle context, const char *name) {
    size_t h = (size_t)-1;
    guarded([&] { h = stream_query((StreamCtx *)context, name); });
    return h;
}

KUN_API const float *kunStreamGetCurrentBuffer(KunStreamContextHandle context, size_t handle) {
    const void *p = nullptr;
    guarded([&] { p = stream_current((StreamCtx *)context, handle); });
    return (const float *)p;
}

KUN_API const void *kunStreamGetCurrentDeviceBuffer(KunStreamContextHandle context, size_t handle) {
    const void *p = nullptr;
    guarded([&] { p = stream_current_device((StreamCtx *)context, handle); });
    return p;
}

KUN_API void kunStreamPushData(KunStreamContextHandle context, size_t handle, const float *buffer) {
    if (guarded([&] { stream_push((StreamCtx *)context, handle, buffer); }))
        fprintf(stderr, "kunStreamPushData: %s\n", last_error());
}

KUN_API void kunStreamRun(KunStreamContextHandle context) {
    if (guarded([&] { stream_run((StreamCtx *)context); })) fprintf(stderr, "kunStreamRun: %s\n", last_error());
}

KUN_API size_t kunStreamTicks(KunStreamContextHandle context) { return stream_ticks((StreamCtx *)context); }

KUN_API void kunDestoryStream(KunStreamContextHandle context) { stream_destroy((StreamCtx *)context); }

// ---- corrWith / aggregrate ----------------------------------------------

KUN_API int kunCorrWith(KunExecutorHandle exec, int layout, int rank_inputs, int num_inputs,
                        const float *const *inputs, const float *corr_with, float *const *outs, size_t num_stocks,
                        size_t total_time, size_t cur_time, size_t length) {
    return guarded([&] {
        if (layout != 0 && layout != 1) throw Error("Unknown layout");
        kunrt::corr_with(**unwrap_exec(exec), false, layout == 0, rank_inputs != 0, num_inputs, inputs, corr_with, outs,
                  num_stocks, total_time, cur_time, length, 8, nullptr);
    });
}

KUN_API int kunCorrWithDevice(KunExecutorHandle exec, int layout, int rank_inputs, int num_inputs,
                              const float *const *inputs, const float *corr_with, float *const *outs,
                              size_t num_stocks, size_t total_time, size_t cur_time, size_t length, void *hip_stream) {
    return guarded([&] {
        if (layout != 0 && layout != 1) throw Error("Unknown layout");
        kunrt::corr_with(**unwrap_exec(exec), true, layout == 0, rank_inputs != 0, num_inputs, inputs, corr_with, outs,
                  num_stocks, total_time, cur_time, length, 8, (hipStream_t)hip_stream);
    });
}

KUN_API int kunAggregrate(KunExecutorHandle exec, int num_aggregrations, const void *const *inputs,
                          const void *const *labels, int dtype, void *const *outs, size_t num_stocks,
                          size_t total_time, size_t cur_time, size_t length) {
    return guarded([&] {
        aggregate(**unwrap_exec(exec), false, num_aggregrations, inputs, labels, dtype, outs, num_stocks, total_time,
                  cur_time, length, 0, nullptr);
    });
}

KUN_API int kunAggregrateDevice(KunExecutorHandle exec, int num_aggregrations, const void *const *inputs,
                                const void *const *labels, int dtype, void *const *outs, size_t out_rows,
                                size_t num_stocks, size_t total_time, size_t cur_time, size_t length,
                                void *hip_stream) {
    return guarded([&] {
        aggregate(**unwrap_exec(exec), true, num_aggregrations, inputs, labels, dtype, outs, num_stocks, total_time,
                  cur_time, length, out_rows, (hipStream_t)hip_stream);
    });
}

// ---- session pools ------------------------------------------------------

KUN_API size_t kunStreamPoolBytesPerStock(KunModuleHandle m) {
    size_t b = 0;
    guarded([&] { b = pool_bytes_per_stock(unwrap_mod(m)); });
    return b;
}

KUN_API KunStreamPoolHandle kunCreateStreamPool(KunExecutorHandle exec, KunModuleHandle m, size_t capacity_stocks,
                                                int max_sessions, int max_inflight) {
    StreamPool *p = nullptr;
    guarded([&] { p = pool_create(*unwrap_exec(exec), unwrap_mod(m), capacity_stocks, max_sessions, max_inflight); });
    return p;
}

KUN_API void kunDestroyStreamPool(KunStreamPoolHandle pool) { pool_destroy((StreamPool *)pool); }

KUN_API size_t kunStreamPoolFreeStocks(KunStreamPoolHandle pool) { return pool_free_stocks((StreamPool *)pool); }

KUN_API int kunStreamPoolAdmit(KunStreamPoolHandle pool, size_t num_stocks, int *session) {
    int sid = -1;
    if (guarded([&] { sid = pool_admit((StreamPool *)pool, num_stocks); })) return 1;
    if (sid < 0) return 2;
    *session = sid;
    return 0;
}

KUN_API int kunStreamPoolRelease(KunStreamPoolHandle pool, int session) {
    return guarded([&] { pool_release((StreamPool *)pool, session); });
}

KUN_API int kunStreamPoolSubmit(KunStreamPoolHandle pool, int nsessions, const int *sessions,
                                const void *const *input_rows, unsigned long long *cycle) {
    return guarded([&] { *cycle = pool_submit((StreamPool *)pool, nsessions, sessions, input_rows); });
}

KUN_API int kunStreamPoolDone(KunStreamPoolHandle pool, unsigned long long cycle, int wait) {
    int r = -1;
    if (guarded([&] { r = pool_done((StreamPool *)pool, cycle, wait != 0); })) return -1;
    return r;
}

KUN_API int kunStreamPoolRead(KunStreamPoolHandle pool, int session, size_t buffer_handle, void *host_row) {
    return guarded([&] { pool_read((StreamPool *)pool, session, buffer_handle, host_row); });
}

KUN_API const void *kunStreamPoolRowDevice(KunStreamPoolHandle pool, int session, size_t buffer_handle) {
    const void *r = nullptr;
    guarded([&] { r = pool_row_device((StreamPool *)pool, session, buffer_handle); });
    return r;
}

KUN_API long kunStreamPoolTicks(KunStreamPoolHandle pool, int session) {
    long t = -1;
    guarded([&] { t = pool_ticks((StreamPool *)pool, session); });
    return t;
}

KUN_API int kunStreamPoolSetReadback(KunStreamPoolHandle pool, int nbuffers, const size_t *buffer_handles) {
    return guarded([&] { pool_set_readback((StreamPool *)pool, nbuffers, buffer_handles); });
}

KUN_API const void *kunStreamPoolResult(KunStreamPoolHandle pool, unsigned long long cycle, int session,
                                        size_t buffer_handle) {
    const void *r = nullptr;
    guarded([&] { r = pool_result((StreamPool *)pool, cycle, session, buffer_handle); });
    return r;
}

KUN_API int kunStreamPoolComputed(KunStreamPoolHandle pool, unsigned long long cycle, int wait) {
    int r = -1;
    if (guarded([&] { r = pool_computed((StreamPool *)pool, cycle, wait != 0); })) return -1;
    return r;
}

KUN_API int kunStreamPoolProfile(KunStreamPoolHandle pool, int enable) {
    return guarded([&] { pool_profile((StreamPool *)pool, enable != 0); });
}

KUN_API int kunStreamPoolPhaseTimes(KunStreamPoolHandle pool, unsigned long long cycle, float *ms, int n) {
    int k = -1;
    if (guarded([&] { k = pool_phase_times((StreamPool *)pool, cycle, ms, n); })) return -1;
    return k;
}

KUN_API int kunStreamPoolAdmitState(KunStreamPoolHandle pool, size_t num_stocks, const char *state, size_t state_len,
                                    int *session) {
    if (!state || !session) {
        set_error("state and session must not be null");
        return 1;
    }
    int sid = -1;
    if (guarded([&] { sid = pool_admit_state((StreamPool *)pool, num_stocks, state, state_len); })) return 1;
    if (sid < 0) return 2;
    *session = sid;
    return 0;
}

KUN_API int kunStreamPoolSerializeStates(KunStreamPoolHandle pool, int session, char *buffer, size_t *size) {
    if (!size || (*size != 0 && !buffer)) {
        set_error("size must not be null, and buffer not null when *size > 0");
        return 1;
    }
    std::string blob;
    if (guarded([&] { blob = pool_serialize((StreamPool *)pool, session); })) return 1;
    const size_t cap = *size;
    *size = blob.size();
    if (cap == 0) return 0;
    if (blob.size() > cap) {
        set_error("buffer too small for the serialised state");
        return 1;
    }
    memcpy(buffer, blob.data(), blob.size());
    return 0;
}

} // extern "C"
