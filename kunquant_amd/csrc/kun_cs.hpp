// Host interface of the cross-sectional row kernels (cross_section.hip).
#pragma once
#include <hip/hip_runtime.h>

namespace kun_cs {

// Where the kernels of a row operation go: launched on a stream, or added
// to a HIP graph under construction (runtime.hip). args: one pointer per
// kernel parameter, as hipLaunchKernel takes them.
struct Sink {
    virtual hipError_t launch(const void *func, dim3 grid, dim3 block, void **args) = 0;
    virtual ~Sink() = default;
};
struct StreamSink : Sink {
    hipStream_t st;
    explicit StreamSink(hipStream_t s) : st(s) {}
    hipError_t launch(const void *func, dim3 grid, dim3 block, void **args) override {
        return hipLaunchKernel(func, grid, block, args, 0, st);
    }
};

// One buffer as seen by a row kernel: element (row, s) lives at
// layout 1 (TS): row * S + s; layout 0 (STs): ((s / blk) * rows + row) * blk + s % blk;
// layout 2 (single row): s. Kernel row t reads row t + row0.
struct View {
    void *ptr;
    int layout;
    long rows;
    long row0;
};

// A contiguous run of n stocks of one row (layout 2), e.g. one session's
// range of a pool row; `in` and `out` point at its first stock.
struct Seg {
    const void *in;
    void *out;
    long n;
};

long max_rank_stocks(int dtype);
// one workgroup per segment; maxS bounds every segment's n
hipError_t launch_rank_segs(int dtype, const Seg *segs, int nseg, long maxS, hipStream_t st);
hipError_t launch_scale_segs(int dtype, const Seg *segs, int nseg, hipStream_t st);
hipError_t launch_rank(int dtype, const View &in, const View &out, long S, long rows, int blk, Sink &sink);
hipError_t launch_rank(int dtype, const View &in, const View &out, long S, long rows, int blk, hipStream_t st);
// row_reduce.hip: per-row correlation (rows start .. start + L - 1 of y; x
// rows from xrow0; sts: STs layout) and per-stock label-group reductions
hipError_t launch_corr(const float *x, int xsts, long xT, long xrow0, const float *y, int ysts, long yT, float *out,
                       long S, long start, long L, int blk, hipStream_t st);
hipError_t launch_aggregate(int dtype, const void *in, const void *labels, void *const *outs, long S, long start,
                            long L, hipStream_t st);
// scratch (scale_scratch_bytes, may be null): row sums of the split form
// (and, float, their reciprocals in double)
size_t scale_scratch_bytes(int dtype, long rows);
hipError_t launch_scale(int dtype, const View &in, const View &out, long S, long rows, int blk, Sink &sink,
                        void *scratch = nullptr);
hipError_t launch_scale(int dtype, const View &in, const View &out, long S, long rows, int blk, hipStream_t st,
                        void *scratch = nullptr);

} // namespace kun_cs
