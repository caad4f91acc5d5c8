// Cross-sectional row kernels: Rank and Scale over the stocks of one row.
//
// Rank (restates cpp/Kun/Rank.hpp:11-57): per row, the non-NaN values are
// sorted; a value's rank is the average 1-based position of its tie group
// [lo, hi): sum = (lo + hi + 1) * (hi - lo) / 2 in integers, then
// T(sum) / T(hi - lo) / T(n). -0 and +0 tie (std::sort / equal_range use
// operator<). On the device a row is sorted with an LDS block radix sort of
// order-preserving integer keys (-0 canonicalised to +0, NaN to the largest
// key) and every element finds its group by two binary searches.
//
// Scale (restates cpp/Kun/Scale.hpp:13-46): sum of |x| over non-NaN values
// accumulated sequentially in stock order (bit-exact order), then
// x / sum, or NaN when x == 0 and sum == 0.
#include <hip/hip_runtime.h>
#include <hipcub/block/block_radix_sort.hpp>
#include <stdint.h>

#include "kun_cs.hpp"
#include "kun_device.hpp"

namespace {

template <typename T> struct KeyOf;
template <> struct KeyOf<float> { using K = uint32_t; };
template <> struct KeyOf<double> { using K = uint64_t; };

template <typename T> __device__ __forceinline__ typename KeyOf<T>::K rank_key(T x) {
    using K = typename KeyOf<T>::K;
    constexpr int NB = sizeof(K) * 8;
    if (x != x) return ~K(0);
    if (x == T(0)) x = T(0);
    K u = kun::to_bits(x);
    return (u >> (NB - 1)) ? ~u : (u | (K(1) << (NB - 1)));
}

__device__ __forceinline__ size_t cs_at(int layout, long row, long s, long S, long rows, int blk) {
    if (layout == 1) return (size_t)row * S + s;
    if (layout == 2) return (size_t)s;
    return ((size_t)(s / blk) * rows + row) * blk + s % blk;
}

template <typename T, int BLOCK, int ITEMS>
__global__ __launch_bounds__(BLOCK) void rank_rows(kun_cs::View in, kun_cs::View out, long S, int blk) {
    using K = typename KeyOf<T>::K;
    using Sort = hipcub::BlockRadixSort<K, BLOCK, ITEMS>;
    __shared__ union {
        typename Sort::TempStorage sort;
        K sorted[BLOCK * ITEMS];
    } sh;
    __shared__ int nvalid;
    const long t = blockIdx.x;
    const T *src = (const T *)in.ptr;
    T *dst = (T *)out.ptr;
    const long rin = t + in.row0;
    if (threadIdx.x == 0) nvalid = 0;
    __syncthreads();
    K keys[ITEMS];
    int cnt = 0;
#pragma unroll
    for (int j = 0; j < ITEMS; j++) {
        long s = threadIdx.x + (long)j * BLOCK;
        K k = ~K(0);
        if (s < S) {
            T x = src[cs_at(in.layout, rin, s, S, in.rows, blk)];
            k = rank_key(x);
            cnt += (x == x) ? 1 : 0;
        }
        keys[j] = k;
    }
    if (cnt) atomicAdd(&nvalid, cnt);
    Sort(sh.sort).Sort(keys);
    __syncthreads();
#pragma unroll
    for (int j = 0; j < ITEMS; j++) sh.sorted[threadIdx.x * ITEMS + j] = keys[j];
    __syncthreads();
    const int n = nvalid;
    for (long s = threadIdx.x; s < S; s += BLOCK) {
        T x = src[cs_at(in.layout, rin, s, S, in.rows, blk)];
        T r;
        if (x != x) {
            r = kun::qnan<T>();
        } else {
            K k = rank_key(x);
            int lo = 0, hi = n;
            while (lo < hi) {
                int mid = (lo + hi) >> 1;
                if (sh.sorted[mid] < k) lo = mid + 1; else hi = mid;
            }
            int a = lo;
            hi = n;
            while (lo < hi) {
                int mid = (lo + hi) >> 1;
                if (k < sh.sorted[mid]) hi = mid; else lo = mid + 1;
            }
            int b = lo;
            int64_t sum = ((int64_t)a + b + 1) * (int64_t)(b - a) / 2;
            r = (T)sum / (T)(b - a) / (T)n;
        }
        dst[cs_at(out.layout, t + out.row0, s, S, out.rows, blk)] = r;
    }
}

constexpr int SCALE_BLOCK = 256;
constexpr int SCALE_CHUNK = 2048;

template <typename T>
__global__ __launch_bounds__(SCALE_BLOCK) void scale_rows(kun_cs::View in, kun_cs::View out, long S, int blk) {
    __shared__ T chunk[SCALE_CHUNK];
    __shared__ T total;
    const long t = blockIdx.x;
    const T *src = (const T *)in.ptr;
    T *dst = (T *)out.ptr;
    const long rin = t + in.row0;
    T sum = T(0);
    for (long c0 = 0; c0 < S; c0 += SCALE_CHUNK) {
        int n = (int)((S - c0) < SCALE_CHUNK ? (S - c0) : SCALE_CHUNK);
        for (int k = threadIdx.x; k < n; k += SCALE_BLOCK) chunk[k] = src[cs_at(in.layout, rin, c0 + k, S, in.rows, blk)];
        __syncthreads();
        if (threadIdx.x == 0) {
            // sequential in stock order, as the reference's scalar loop
#pragma unroll 8
            for (int k = 0; k < n; k++) {
                T x = chunk[k];
                if (x == x) sum = sum + kun::vabs(x);
            }
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) total = sum;
    __syncthreads();
    sum = total;
    for (long s = threadIdx.x; s < S; s += SCALE_BLOCK) {
        T x = src[cs_at(in.layout, rin, s, S, in.rows, blk)];
        dst[cs_at(out.layout, t + out.row0, s, S, out.rows, blk)] =
            (x == T(0) && sum == T(0)) ? kun::qnan<T>() : x / sum;
    }
}

template <typename T, int BLOCK, int ITEMS>
hipError_t launch_rank_n(const kun_cs::View &in, const kun_cs::View &out, long S, long rows, int blk,
                         hipStream_t st) {
    hipLaunchKernelGGL((rank_rows<T, BLOCK, ITEMS>), dim3((unsigned)rows), dim3(BLOCK), 0, st, in, out, S, blk);
    return hipGetLastError();
}

template <typename T>
hipError_t launch_rank_t(const kun_cs::View &in, const kun_cs::View &out, long S, long rows, int blk,
                         hipStream_t st) {
    if (S <= 256) return launch_rank_n<T, 256, 1>(in, out, S, rows, blk, st);
    if (S <= 512) return launch_rank_n<T, 256, 2>(in, out, S, rows, blk, st);
    if (S <= 1024) return launch_rank_n<T, 256, 4>(in, out, S, rows, blk, st);
    if (S <= 2048) return launch_rank_n<T, 256, 8>(in, out, S, rows, blk, st);
    if (S <= 4096) return launch_rank_n<T, 256, 16>(in, out, S, rows, blk, st);
    if (S <= 8192) return launch_rank_n<T, 256, 32>(in, out, S, rows, blk, st);
    if (S <= 16384) return launch_rank_n<T, 512, 32>(in, out, S, rows, blk, st);
    return hipErrorInvalidValue;
}

} // namespace

namespace kun_cs {

long max_rank_stocks(int dtype) { return 16384; }

hipError_t launch_rank(int dtype, const View &in, const View &out, long S, long rows, int blk, hipStream_t st) {
    if (rows <= 0) return hipSuccess;
    return dtype ? launch_rank_t<double>(in, out, S, rows, blk, st) : launch_rank_t<float>(in, out, S, rows, blk, st);
}

hipError_t launch_scale(int dtype, const View &in, const View &out, long S, long rows, int blk, hipStream_t st) {
    if (rows <= 0) return hipSuccess;
    if (dtype)
        hipLaunchKernelGGL(scale_rows<double>, dim3((unsigned)rows), dim3(SCALE_BLOCK), 0, st, in, out, S, blk);
    else
        hipLaunchKernelGGL(scale_rows<float>, dim3((unsigned)rows), dim3(SCALE_BLOCK), 0, st, in, out, S, blk);
    return hipGetLastError();
}

} // namespace kun_cs
