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
// chunk of every row-chunked component (0: one chunk). Components run in
// descriptor order (the device runtime orders them by cost; results do not
// depend on the order).
extern "C" int kun_emu_run_stage(const KunModuleDesc *m, int stage, void **bufs, long S, long T, long start,
                                 long L, long chunk_rows) {
    const KunStageDesc &sd = m->stages[stage];
    if (sd.kind != KUN_STAGE_TS || L <= 0) return 1;
    const size_t esz = m->dtype == KUN_DTYPE_DOUBLE ? 8 : 4;
    std::vector<char> mat((size_t)sd.nmat * L * S * esz + 1);
    const long nsb = (S + 63) / 64;
    for (int kj = 0; kj < sd.nkern; kj++) {
        const KunKernelDesc &kd = sd.kerns[kj];
        if (kd.ncomp > KUN_MAX_COMPS) return 2;
        kun::StageArgs a;
        memset(&a, 0, sizeof a);
        int n = 0;
        for (int j = 0; j < sd.nin; j++) a.bufs[n++] = bufs[sd.in[j]];
        for (int j = 0; j < sd.nout; j++) a.bufs[n++] = bufs[sd.out[j]];
        for (int j = 0; j < sd.nmat; j++) a.bufs[n++] = mat.data() + (size_t)j * L * S * esz;
        a.S = S;
        a.T = T;
        a.start = start;
        a.L = L;
        a.blk = m->blocking;
        a.nz = kd.ncomp;
        size_t off = 0;
        long blocks = 0;
        for (int z = 0; z < kd.ncomp; z++) {
            long chunk = (kd.kind == KUN_KERNEL_PAR && chunk_rows > 0) ? chunk_rows : L;
            long nch = (L + chunk - 1) / chunk;
            a.zcomp[z] = z;
            a.zfirst[z] = (int)blocks;
            a.zchunk[z] = (int)chunk;
            a.znch[z] = (int)nch;
            a.zscr[z] = (long)off;
            off += (size_t)kd.comp_spill[z] * S * nch;
            blocks += nsb * nch;
        }
        a.zfirst[kd.ncomp] = (int)blocks;
        std::vector<char> scratch(off + 1);
        a.scratch = scratch.data();
        gridDim = {(unsigned)blocks, 1, 1};
        blockDim = {(unsigned)sd.block, 1, 1};
        for (unsigned x = 0; x < gridDim.x; x++)
            for (unsigned t = 0; t < blockDim.x; t++) {
                blockIdx = {x, 0, 0};
                threadIdx = {t, 0, 0};
                ((StageKernel)kd.kernel)(a);
            }
    }
    return 0;
}
