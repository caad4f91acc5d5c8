// Test-only CPU driver for generated batch stage kernels: launches every
// kernel of one time-series stage over its grid (workgroups, chunks,
// components) one lane at a time, with the scratch layout the device runtime
// uses (kunquant_amd/csrc/runtime.hip, batch_layout / enqueue_batch).
#include <hip/hip_runtime.h>

#include <stdlib.h>
#include <string.h>

#include <vector>

#include "kun_device.hpp"
#include "kun_module_abi.h"

thread_local kun_emu_dim3 blockIdx, threadIdx, gridDim, blockDim;

typedef void (*StageKernel)(kun::StageArgs);

// bufs: device-layout pointers of every module buffer; chunk_rows: rows per
// chunk of every row-chunked kernel (0: one chunk).
extern "C" int kun_emu_run_stage(const KunModuleDesc *m, int stage, void **bufs, long S, long T, long start,
                                 long L, long chunk_rows) {
    const KunStageDesc &sd = m->stages[stage];
    if (sd.kind != KUN_STAGE_TS || L <= 0) return 1;
    const size_t esz = m->dtype == KUN_DTYPE_DOUBLE ? 8 : 4;
    std::vector<char> mat((size_t)sd.nmat * L * S * esz + 1);
    for (int kj = 0; kj < sd.nkern; kj++) {
        const KunKernelDesc &kd = sd.kerns[kj];
        long chunk = (kd.kind == KUN_KERNEL_PAR && chunk_rows > 0) ? chunk_rows : L;
        long nch = (L + chunk - 1) / chunk;
        std::vector<char> scratch((size_t)kd.spill_bytes * S * nch + 1);
        kun::StageArgs a;
        memset(&a, 0, sizeof a);
        int n = 0;
        for (int j = 0; j < sd.nin; j++) a.bufs[n++] = bufs[sd.in[j]];
        for (int j = 0; j < sd.nout; j++) a.bufs[n++] = bufs[sd.out[j]];
        a.S = S;
        a.T = T;
        a.start = start;
        a.L = L;
        a.chunk = chunk;
        a.blk = m->blocking;
        a.scratch = scratch.data();
        a.mat = mat.data();
        gridDim = {(unsigned)((S + sd.block - 1) / sd.block), (unsigned)nch, (unsigned)kd.ncomp};
        blockDim = {(unsigned)sd.block, 1, 1};
        for (unsigned z = 0; z < gridDim.z; z++)
            for (unsigned y = 0; y < gridDim.y; y++)
                for (unsigned x = 0; x < gridDim.x; x++)
                    for (unsigned t = 0; t < blockDim.x; t++) {
                        blockIdx = {x, y, z};
                        threadIdx = {t, 0, 0};
                        ((StageKernel)kd.kernel)(a);
                    }
    }
    return 0;
}
