// Batch runs on device-resident buffers that allocate no device memory:
// the run's arena (module temporaries, values handed between a stage's
// kernels, row-chunk spill columns, Scale row sums) lives in one file-scope
// device array instead of memory from hipMalloc. For callers that cannot
// allocate (code under HIP graph capture, tight serving loops).
//
// The array is each device's own copy of a module-scope variable of this
// library's code object: its device address comes from hipGetSymbolAddress,
// looked up once per device on the first call made on it, and only that
// address is handed to kernels (the variable's name in host code is a
// host-side stand-in, not a device address).
//
// Every call on a device shares the array and leaves the last call's
// contents in it; no kernel reads arena memory before the same run writes
// it (history rings are NaN-filled by their kernels, row sums are written
// before they are read), so nothing is cleared between calls. Calls that
// overlap on one device race on the array: the caller orders them (one
// stream, or events).
#include <hip/hip_runtime.h>

#include <mutex>
#include <string>
#include <vector>

#include "../../include/kun_capi.h"
#include "kun_rt.hpp"

namespace kunrt {

static __device__ __align__(256) char scratch[KUN_NOALLOC_SCRATCH_BYTES];

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
char *g_scratch_ptr[kMaxDevices] = {};

// device address of `scratch` on `device` (current device), looked up once
char *scratch_on(int device) {
    if (device < 0 || device >= kMaxDevices) throw Error("device index out of range for the no-allocation arena");
    std::lock_guard<std::mutex> g(g_scratch_mu);
    if (!g_scratch_ptr[device]) {
        void *ptr = nullptr;
        hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !ptr)
            throw Error(std::string("hipGetSymbolAddress of the no-allocation arena failed: ") + hipGetErrorString(e));
        g_scratch_ptr[device] = (char *)ptr;
    }
    return g_scratch_ptr[device];
}
} // namespace

void run_graph_device_noalloc(Executor &ex, const KunModuleDesc *m, const NameMap &bufs, size_t S, size_t T,
                              size_t start, size_t L, hipStream_t st) {
    check_batch_module(m, S, T, start, L);
    std::lock_guard<std::mutex> g(ex.mu);
    HIPCHECK(hipSetDevice(ex.device));
    if (!st) st = ex.stream;
    const BatchLayout lay = batch_layout(m, S, L, ex.policy_for(m, S, L), (int)ex.lanes.size());
    if (lay.bytes > (size_t)KUN_NOALLOC_SCRATCH_BYTES)
        throw Error("kunRunGraphDeviceNoAlloc: this run needs " + std::to_string(lay.bytes) +
                    " bytes of device scratch, more than the " + std::to_string((size_t)KUN_NOALLOC_SCRATCH_BYTES) +
                    " bytes of its fixed arena (use kunRunGraphDevice)");
    char *arena = scratch_on(ex.device);
    std::vector<void *> dev(m->nbuf, nullptr);
    for (int b = 0; b < m->nbuf; b++) {
        const KunBufferDesc &d = m->bufs[b];
        if (d.kind == KUN_BUF_TEMP) {
            dev[b] = arena + lay.buf_off[b];
        } else {
            auto f = bufs.find(d.name);
            if (f == bufs.end()) throw Error(std::string("Buffer name not found: ") + d.name);
            dev[b] = f->second;
        }
    }
    // every kernel in launch order on the caller's stream: no lanes, so no
    // events to create either
    enqueue_batch_serial(m, dev, lay, arena, S, T, start, L, st);
}

} // namespace kunrt

using namespace kunrt;

extern "C" KUN_API int kunRunGraphDeviceNoAlloc(KunExecutorHandle exec, KunModuleHandle m,
                                                KunBufferNameMapHandle buffers, size_t num_stocks, size_t total_time,
                                                size_t cur_time, size_t length, void *hip_stream) {
    try {
        set_error("");
        run_graph_device_noalloc(**(std::shared_ptr<Executor> *)exec, (const KunModuleDesc *)m,
                                 *(const NameMap *)buffers, num_stocks, total_time, cur_time, length,
                                 (hipStream_t)hip_stream);
        return 0;
    } catch (const std::exception &e) {
        set_error(e.what());
        return 1;
    }
}
