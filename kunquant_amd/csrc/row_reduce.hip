// Row and column reductions after the factors: corrWith (per-day
// correlation of each factor with a target, e.g. forward returns) and
// aggregrate (per-stock reductions over runs of equal labels, e.g. bars into
// days).
//
// corrWith restates cpp/Kun/CorrWith.hpp:11-121: per row, sums over the
// stocks in stock order of x, y, x*x, y*y and x*y (no NaN skipping), each
// divided by the stock count, then (XY - X*Y) / (sqrt(X2 - X*X) *
// sqrt(Y2 - Y*Y)). The five sums are sequential chains in stock order (bit
// parity); one lane owns one row and walks its stocks, the five chains
// interleaved, loads issued a group of stocks ahead. With rank inputs, x is
// the row's rank, computed by the Rank row kernel first.
//
// aggregrate restates cpp/Kun/Aggregration.cpp:11-105: one lane per stock
// walks the rows in order; a change of label closes a group and stores its
// sum (kahanAdd), min and max (NaN if any NaN), first, last, count (every
// row) and mean (sum / count) at the group's index.
#include <hip/hip_runtime.h>

#include "kun_cs.hpp"
#include "kun_device.hpp"

namespace {

constexpr int CORR_UNROLL = 8;

__device__ __forceinline__ size_t rr_at(int sts, long t, long s, long S, long T, int blk) {
    return sts ? ((size_t)(s / blk) * T + t) * blk + s % blk : (size_t)t * S + s;
}

// rows [start, start + L) of x (layout `xsts`, rows of T) and y (`ysts`);
// x may instead be a [L][S] TS array of ranks (xrows = L, xrow0 = 0)
__global__ __launch_bounds__(64) void corr_rows(const float *x, int xsts, long xT, long xrow0, const float *y,
                                                int ysts, long yT, float *out, long S, long start, long L, int blk) {
    const long r = (long)blockIdx.x * 64 + threadIdx.x;
    if (r >= L) return;
    const long tx = r + xrow0, ty = r + start;
    float XY = 0.f, X = 0.f, Y = 0.f, X2 = 0.f, Y2 = 0.f;
    long s = 0;
    for (; s + CORR_UNROLL <= S; s += CORR_UNROLL) {
        float a[CORR_UNROLL], b[CORR_UNROLL];
#pragma unroll
        for (int k = 0; k < CORR_UNROLL; k++) {
            a[k] = x[rr_at(xsts, tx, s + k, S, xT, blk)];
            b[k] = y[rr_at(ysts, ty, s + k, S, yT, blk)];
        }
#pragma unroll
        for (int k = 0; k < CORR_UNROLL; k++) {
            X = X + a[k];
            Y = Y + b[k];
            X2 = X2 + a[k] * a[k];
            Y2 = Y2 + b[k] * b[k];
            XY = XY + a[k] * b[k];
        }
    }
    for (; s < S; s++) {
        const float a = x[rr_at(xsts, tx, s, S, xT, blk)], b = y[rr_at(ysts, ty, s, S, yT, blk)];
        X = X + a;
        Y = Y + b;
        X2 = X2 + a * a;
        Y2 = Y2 + b * b;
        XY = XY + a * b;
    }
    const float n = (float)S;
    XY = XY / n;
    X = X / n;
    Y = Y / n;
    X2 = X2 / n;
    Y2 = Y2 / n;
    out[r] = (XY - X * Y) / (sqrtf(X2 - X * X) * sqrtf(Y2 - Y * Y));
}

struct AggOut {
    void *p[7]; // sum, min, max, first, last, count, mean (null: not wanted)
};

template <typename T>
__global__ __launch_bounds__(64) void aggregate_cols(const T *in, const T *labels, AggOut o, long S, long start,
                                                     long L) {
    const long s = (long)blockIdx.x * 64 + threadIdx.x;
    if (s >= S || L <= 0) return;
    const T inf = kun::inf<T>();
    T sum = 0, comp = 0, mn = inf, mx = -inf, last = 0, count = 0;
    T first = in[(size_t)start * S + s];
    T last_label = labels[start];
    long g = 0;
    auto flush = [&]() {
        const T vals[7] = {sum, mn, mx, first, last, count, sum / count};
#pragma unroll
        for (int k = 0; k < 7; k++)
            if (o.p[k]) ((T *)o.p[k])[(size_t)g * S + s] = vals[k];
    };
    for (long i = 0; i < L; i++) {
        const T label = labels[start + i];
        const T cur = in[(size_t)(start + i) * S + s];
        if (label != last_label) {
            flush();
            g++;
            first = cur;
            sum = 0;
            comp = 0;
            mn = inf;
            mx = -inf;
            count = 0;
        }
        count = count + T(1);
        last = cur;
        last_label = label;
        mx = (mx != mx || cur != cur) ? kun::qnan<T>() : (mx > cur ? mx : cur);
        mn = (mn != mn || cur != cur) ? kun::qnan<T>() : (mn < cur ? mn : cur);
        const T y = cur - comp;
        const T t = sum + y;
        comp = (t - sum) - y;
        sum = t;
    }
    flush();
}

} // namespace

namespace kun_cs {

hipError_t launch_corr(const float *x, int xsts, long xT, long xrow0, const float *y, int ysts, long yT, float *out,
                       long S, long start, long L, int blk, hipStream_t st) {
    if (L <= 0) return hipSuccess;
    hipLaunchKernelGGL(corr_rows, dim3((unsigned)((L + 63) / 64)), dim3(64), 0, st, x, xsts, xT, xrow0, y, ysts, yT,
                       out, S, start, L, blk);
    return hipGetLastError();
}

hipError_t launch_aggregate(int dtype, const void *in, const void *labels, void *const *outs, long S, long start,
                            long L, hipStream_t st) {
    if (L <= 0 || S <= 0) return hipSuccess;
    AggOut o;
    for (int k = 0; k < 7; k++) o.p[k] = outs[k];
    const dim3 grid((unsigned)((S + 63) / 64));
    if (dtype)
        hipLaunchKernelGGL(aggregate_cols<double>, grid, dim3(64), 0, st, (const double *)in, (const double *)labels, o,
                           S, start, L);
    else
        hipLaunchKernelGGL(aggregate_cols<float>, grid, dim3(64), 0, st, (const float *)in, (const float *)labels, o, S,
                           start, L);
    return hipGetLastError();
}

} // namespace kun_cs

// ---- host drivers -----------------------------------------------------------
#include <string.h>

#include <string>
#include <vector>

#include "kun_rt.hpp"

namespace kunrt {

namespace {

struct DevBuf {
    Executor &ex;
    void *p = nullptr;
    size_t n = 0;
    DevBuf(Executor &e, size_t bytes) : ex(e), p(e.pool.get(bytes)), n(bytes) {}
    ~DevBuf() { ex.pool.put(p, n); }
};

size_t count_groups(const char *labels, size_t esz, size_t start, size_t L) {
    if (L == 0) return 0;
    size_t g = 1;
    for (size_t i = start + 1; i < start + L; i++) {
        const bool diff = esz == 8 ? ((const double *)labels)[i] != ((const double *)labels)[i - 1]
                                   : ((const float *)labels)[i] != ((const float *)labels)[i - 1];
        g += diff ? 1 : 0;
    }
    return g;
}

} // namespace

void corr_with(Executor &ex, bool device, int sts, bool rank_inputs, int n, const float *const *inputs,
               const float *target, float *const *outs, size_t S, size_t T, size_t start, size_t L, int blk,
               hipStream_t st) {
    if (start + L > T) throw Error("Bad parameter: cur_time + length exceeds total_time");
    if (n < 0) throw Error("bad number of inputs");
    if (rank_inputs && (long)S > kun_cs::max_rank_stocks(0))
        throw Error("Rank over more than " + std::to_string(kun_cs::max_rank_stocks(0)) + " stocks is not supported");
    std::lock_guard<std::mutex> g(ex.mu);
    HIPCHECK(hipSetDevice(ex.device));
    if (!st) st = ex.stream;
    const size_t in_bytes = (sts ? (S + blk - 1) / blk * blk : S) * T * sizeof(float);
    std::unique_ptr<DevBuf> dtarget, dx, dout;
    const float *y = target;
    if (!device) {
        dtarget.reset(new DevBuf(ex, in_bytes));
        HIPCHECK(hipMemcpyAsync(dtarget->p, target, in_bytes, hipMemcpyHostToDevice, st));
        y = (const float *)dtarget->p;
        dx.reset(new DevBuf(ex, in_bytes));
        dout.reset(new DevBuf(ex, L * sizeof(float) + 4));
    }
    std::unique_ptr<DevBuf> ranks;
    if (rank_inputs) ranks.reset(new DevBuf(ex, S * L * sizeof(float) + 4));
    for (int i = 0; i < n; i++) {
        const float *x = inputs[i];
        float *o = outs[i];
        if (!device) {
            HIPCHECK(hipMemcpyAsync(dx->p, x, in_bytes, hipMemcpyHostToDevice, st));
            x = (const float *)dx->p;
            o = (float *)dout->p;
        }
        hipError_t e;
        if (rank_inputs) {
            kun_cs::View vi{(void *)x, sts ? 0 : 1, (long)T, (long)start}, vo{ranks->p, 1, (long)L, 0};
            e = kun_cs::launch_rank(0, vi, vo, (long)S, (long)L, blk, st);
            if (e == hipSuccess)
                e = kun_cs::launch_corr((const float *)ranks->p, 0, (long)L, 0, y, sts, (long)T, o, (long)S,
                                        (long)start, (long)L, blk, st);
        } else {
            e = kun_cs::launch_corr(x, sts, (long)T, (long)start, y, sts, (long)T, o, (long)S, (long)start, (long)L,
                                    blk, st);
        }
        if (e != hipSuccess) throw Error(std::string("corrWith kernel: ") + hipGetErrorString(e));
        if (!device) HIPCHECK(hipMemcpyAsync(outs[i], o, L * sizeof(float), hipMemcpyDeviceToHost, st));
    }
    // host buffers are copied back; pooled scratch returns to the pool only
    // once the kernels reading it are done
    if (!device || ranks) HIPCHECK(hipStreamSynchronize(st));
}

void aggregate(Executor &ex, bool device, int n, const void *const *inputs, const void *const *labels, int dtype,
               void *const *outs, size_t S, size_t T, size_t start, size_t L, size_t out_rows, hipStream_t st) {
    if (start + L > T) throw Error("Bad parameter: cur_time + length exceeds total_time");
    if (n < 0) throw Error("bad number of aggregrations");
    const size_t esz = dtype ? 8 : 4;
    std::lock_guard<std::mutex> g(ex.mu);
    HIPCHECK(hipSetDevice(ex.device));
    if (!st) st = ex.stream;
    std::vector<char> hl;
    for (int i = 0; i < n; i++) {
        // group count from the labels (copied back when they live on the device)
        const char *lab = (const char *)labels[i];
        if (device) {
            hl.resize(T * esz);
            HIPCHECK(hipMemcpyAsync(hl.data(), labels[i], T * esz, hipMemcpyDeviceToHost, st));
            HIPCHECK(hipStreamSynchronize(st));
            lab = hl.data();
        }
        const size_t groups = count_groups(lab, esz, start, L);
        if (device && groups > out_rows)
            throw Error("aggregrate: " + std::to_string(groups) + " label groups exceed the " +
                        std::to_string(out_rows) + " output rows");
        void *dev_outs[7];
        std::unique_ptr<DevBuf> din, dlab, douts[7];
        const void *x = inputs[i], *lb = labels[i];
        if (!device) {
            din.reset(new DevBuf(ex, S * T * esz + 8));
            dlab.reset(new DevBuf(ex, T * esz + 8));
            HIPCHECK(hipMemcpyAsync(din->p, inputs[i], S * T * esz, hipMemcpyHostToDevice, st));
            HIPCHECK(hipMemcpyAsync(dlab->p, labels[i], T * esz, hipMemcpyHostToDevice, st));
            x = din->p;
            lb = dlab->p;
        }
        for (int k = 0; k < 7; k++) {
            void *o = outs[(size_t)i * 7 + k];
            dev_outs[k] = o;
            if (o && !device) {
                douts[k].reset(new DevBuf(ex, groups * S * esz + 8));
                dev_outs[k] = douts[k]->p;
            }
        }
        hipError_t e = kun_cs::launch_aggregate(dtype, x, lb, dev_outs, (long)S, (long)start, (long)L, st);
        if (e != hipSuccess) throw Error(std::string("aggregrate kernel: ") + hipGetErrorString(e));
        if (!device) {
            for (int k = 0; k < 7; k++)
                if (outs[(size_t)i * 7 + k])
                    HIPCHECK(hipMemcpyAsync(outs[(size_t)i * 7 + k], dev_outs[k], groups * S * esz,
                                            hipMemcpyDeviceToHost, st));
            HIPCHECK(hipStreamSynchronize(st));
        }
    }
}

} // namespace kunrt
