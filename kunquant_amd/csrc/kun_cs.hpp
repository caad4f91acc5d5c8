// Host interface of the cross-sectional row kernels (cross_section.hip).
#pragma once
#include <hip/hip_runtime.h>

namespace kun_cs {

// One buffer as seen by a row kernel: element (row, s) lives at
// layout 1 (TS): row * S + s; layout 0 (STs): ((s / blk) * rows + row) * blk + s % blk;
// layout 2 (single row): s. Kernel row t reads row t + row0.
struct View {
    void *ptr;
    int layout;
    long rows;
    long row0;
};

long max_rank_stocks(int dtype);
hipError_t launch_rank(int dtype, const View &in, const View &out, long S, long rows, int blk, hipStream_t st);
hipError_t launch_scale(int dtype, const View &in, const View &out, long S, long rows, int blk, hipStream_t st);

} // namespace kun_cs
