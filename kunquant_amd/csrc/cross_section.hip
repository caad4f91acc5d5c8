// Cross-sectional row kernels: Rank and Scale over the stocks of one row.
//
// Rank (restates cpp/Kun/Rank.hpp:11-57): per row, the non-NaN values are
// sorted; a value's rank is the average 1-based position of its tie group
// [lo, hi): sum = (lo + hi + 1) * (hi - lo) / 2 in integers, then
// T(sum) / T(hi - lo) / T(n). -0 and +0 tie (std::sort / equal_range use
// operator<). On the device a row's (key, stock) pairs are sorted with an
// LDS block radix sort of order-preserving integer keys (-0 canonicalised to
// +0, NaN to the largest key) and tie groups are read off the sorted order.
//
// Scale (restates cpp/Kun/Scale.hpp:13-46): sum of |x| over non-NaN values
// accumulated sequentially in stock order (bit-exact order), then
// x / sum, or NaN when x == 0 and sum == 0.
#include <hip/hip_runtime.h>
#include <rocprim/block/block_radix_sort.hpp>
#include <rocprim/block/block_scan.hpp>
#include <rocprim/block/block_sort.hpp>
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

// One workgroup per row. (key, stock) pairs are radix sorted; in the sorted
// order the tie group [a, b) of position p is found without searching:
// a = the last group start at or before p (a block max-scan of group
// starts), b = the end recorded by the group's last element. Ranks go to LDS
// at their stock index and leave in one coalesced pass.
template <typename K, int BLOCK, int ITEMS, bool MERGE> struct PairSort;
template <typename K, int BLOCK, int ITEMS> struct PairSort<K, BLOCK, ITEMS, false> {
    using Impl = rocprim::block_radix_sort<K, BLOCK, ITEMS, int>;
    using storage_type = typename Impl::storage_type;
    __device__ __forceinline__ void sort(K (&k)[ITEMS], int (&v)[ITEMS], storage_type &st) { Impl().sort(k, v, st); }
};
template <typename K, int BLOCK, int ITEMS> struct PairSort<K, BLOCK, ITEMS, true> {
    using Impl = rocprim::block_sort<K, BLOCK, ITEMS, int, rocprim::block_sort_algorithm::merge_sort>;
    using storage_type = typename Impl::storage_type;
    __device__ __forceinline__ void sort(K (&k)[ITEMS], int (&v)[ITEMS], storage_type &st) {
        Impl().sort(k, v, st, [](const K &a, const K &b) { return a < b; });
    }
};

// MERGE: rocPRIM's merge sort (faster than its radix sort for large rows);
// either sort serves, since ranks only need every element inside its tie
// group.
template <typename T, int BLOCK, int ITEMS, bool MERGE>
__device__ __forceinline__ void rank_row(const kun_cs::View &in, const kun_cs::View &out, long S, int blk,
                                         const long t) {
    using K = typename KeyOf<T>::K;
    constexpr int N = BLOCK * ITEMS;
    using Sort = PairSort<K, BLOCK, ITEMS, MERGE>;
    using Scan = rocprim::block_scan<int, BLOCK>;
    static_assert(sizeof(T) <= sizeof(K) && sizeof(int) <= sizeof(K), "LDS aliasing");
    __shared__ union {
        typename Sort::storage_type sort;
        K skey[N];  // sorted keys; then group ends (int); then ranks by stock (T)
    } sh;
    __shared__ typename Scan::storage_type scan;
    __shared__ int nvalid;
    int *endof = (int *)sh.skey;
    T *res = (T *)sh.skey;
    const T *src = (const T *)in.ptr;
    T *dst = (T *)out.ptr;
    const long rin = t + in.row0;
    if (threadIdx.x == 0) nvalid = 0;
    __syncthreads();
    K keys[ITEMS];
    int idx[ITEMS];
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
        idx[j] = (int)s;
    }
    if (cnt) atomicAdd(&nvalid, cnt);
    Sort().sort(keys, idx, sh.sort);
    __syncthreads();
    const int p0 = threadIdx.x * ITEMS;  // blocked: this thread holds sorted positions p0 .. p0+ITEMS-1
#pragma unroll
    for (int j = 0; j < ITEMS; j++) sh.skey[p0 + j] = keys[j];
    __syncthreads();
    const int n = nvalid;
    int start[ITEMS];
    bool last[ITEMS];
    int run = -1;
#pragma unroll
    for (int j = 0; j < ITEMS; j++) {
        const int p = p0 + j;
        if (p == 0 || sh.skey[p - 1] != keys[j]) run = p;
        start[j] = run;
        last[j] = p < n && (p == n - 1 || sh.skey[p + 1] != keys[j]);
    }
    int prefix;
    Scan().exclusive_scan(run, prefix, -1, scan, rocprim::maximum<int>());
    __syncthreads();  // every read of the sorted keys is done; reuse their LDS
#pragma unroll
    for (int j = 0; j < ITEMS; j++) {
        start[j] = start[j] > prefix ? start[j] : prefix;
        if (last[j]) endof[start[j]] = p0 + j + 1;
    }
    __syncthreads();
    T r[ITEMS];
#pragma unroll
    for (int j = 0; j < ITEMS; j++) {
        const int p = p0 + j;
        if (p < n) {
            const int a = start[j], b = endof[a];
            const int64_t sum = ((int64_t)a + b + 1) * (int64_t)(b - a) / 2;
            // x / 1 == x exactly: a group of one skips the first division
            const T q = (b - a == 1) ? (T)sum : (T)sum / (T)(b - a);
            r[j] = q / (T)n;
        } else {
            r[j] = kun::qnan<T>();
        }
    }
    __syncthreads();
#pragma unroll
    for (int j = 0; j < ITEMS; j++)
        if (idx[j] < S) res[idx[j]] = r[j];
    __syncthreads();
    for (long s = threadIdx.x; s < S; s += BLOCK) dst[cs_at(out.layout, t + out.row0, s, S, out.rows, blk)] = res[s];
}

// Keys-only form: the keys alone are merge sorted (no stock payload), the
// sorted keys go to LDS, and every element finds its tie group [a, b) by
// binary search -- a = first sorted position holding its key, b = one past
// the last -- then leaves straight from its registers, coalesced. a and b are
// the group bounds of the pair form, so the ranks are the same numbers.
template <typename T, int BLOCK, int ITEMS>
__device__ __forceinline__ void rank_row_keys(const kun_cs::View &in, const kun_cs::View &out, long S, int blk,
                                              const long t) {
    using K = typename KeyOf<T>::K;
    constexpr int N = BLOCK * ITEMS;
    using Sort = rocprim::block_sort<K, BLOCK, ITEMS, rocprim::empty_type, rocprim::block_sort_algorithm::merge_sort>;
    __shared__ union {
        typename Sort::storage_type sort;
        K skey[N];
    } sh;
    __shared__ int nvalid;
    const T *src = (const T *)in.ptr;
    T *dst = (T *)out.ptr;
    const long rin = t + in.row0;
    if (threadIdx.x == 0) nvalid = 0;
    __syncthreads();
    K keys[ITEMS], mine[ITEMS];
    int cnt = 0;
#pragma unroll
    for (int j = 0; j < ITEMS; j++) {
        const long s = threadIdx.x + (long)j * BLOCK;
        K k = ~K(0);
        if (s < S) {
            T x = src[cs_at(in.layout, rin, s, S, in.rows, blk)];
            k = rank_key(x);
            cnt += (x == x) ? 1 : 0;
        }
        keys[j] = k;
        mine[j] = k;
    }
    if (cnt) atomicAdd(&nvalid, cnt);
    Sort().sort(keys, sh.sort, [](const K &a, const K &b) { return a < b; });
    __syncthreads();
    const int p0 = threadIdx.x * ITEMS;
#pragma unroll
    for (int j = 0; j < ITEMS; j++) sh.skey[p0 + j] = keys[j];
    __syncthreads();
    const int n = nvalid;
    // lower bounds of every item, the ITEMS searches interleaved
    int lo[ITEMS], len[ITEMS];
#pragma unroll
    for (int j = 0; j < ITEMS; j++) {
        lo[j] = 0;
        len[j] = n;
    }
    for (int step = 0; step < 32; step++) {
        bool any = false;
#pragma unroll
        for (int j = 0; j < ITEMS; j++) {
            if (len[j] > 0) {
                const int half = len[j] >> 1;
                if (sh.skey[lo[j] + half] < mine[j]) {
                    lo[j] += half + 1;
                    len[j] -= half + 1;
                } else {
                    len[j] = half;
                }
                any = true;
            }
        }
        if (!any) break;
    }
    // upper bounds, searched only for elements with ties
    int hi[ITEMS];
#pragma unroll
    for (int j = 0; j < ITEMS; j++) {
        hi[j] = lo[j] + 1;
        len[j] = (mine[j] != ~K(0) && hi[j] < n && sh.skey[hi[j]] == mine[j]) ? n - hi[j] : 0;
    }
    for (int step = 0; step < 32; step++) {
        bool any = false;
#pragma unroll
        for (int j = 0; j < ITEMS; j++) {
            if (len[j] > 0) {
                const int half = len[j] >> 1;
                if (!(mine[j] < sh.skey[hi[j] + half])) {
                    hi[j] += half + 1;
                    len[j] -= half + 1;
                } else {
                    len[j] = half;
                }
                any = true;
            }
        }
        if (!any) break;
    }
#pragma unroll
    for (int j = 0; j < ITEMS; j++) {
        const long s = threadIdx.x + (long)j * BLOCK;
        if (s >= S) continue;
        T r = kun::qnan<T>();
        if (mine[j] != ~K(0) && lo[j] < n) {
            const int a = lo[j], b = hi[j];
            const int64_t sum = ((int64_t)a + b + 1) * (int64_t)(b - a) / 2;
            const T q = (b - a == 1) ? (T)sum : (T)sum / (T)(b - a);
            r = q / (T)n;
        }
        dst[cs_at(out.layout, t + out.row0, s, S, out.rows, blk)] = r;
    }
}

template <typename T, int BLOCK, int ITEMS>
__global__ __launch_bounds__(BLOCK) void rank_rows_keys(kun_cs::View in, kun_cs::View out, long S, int blk) {
    rank_row_keys<T, BLOCK, ITEMS>(in, out, S, blk, blockIdx.x);
}

// one workgroup per row of the views
template <typename T, int BLOCK, int ITEMS, bool MERGE>
__global__ __launch_bounds__(BLOCK) void rank_rows(kun_cs::View in, kun_cs::View out, long S, int blk) {
    rank_row<T, BLOCK, ITEMS, MERGE>(in, out, S, blk, blockIdx.x);
}

// one workgroup per segment (a contiguous run of stocks of one row)
template <typename T, int BLOCK, int ITEMS, bool MERGE>
__global__ __launch_bounds__(BLOCK) void rank_segs(const kun_cs::Seg *segs) {
    const kun_cs::Seg g = segs[blockIdx.x];
    const kun_cs::View in{const_cast<void *>(g.in), 2, 1, 0}, out{g.out, 2, 1, 0};
    rank_row<T, BLOCK, ITEMS, MERGE>(in, out, g.n, 1, 0);
}

// ROWS rows per workgroup; the sum of each row is one sequential chain (lane
// r of the first wave owns row r), fed from LDS tiles that the whole
// workgroup loads one tile ahead and stores already masked: |x|, or +0 for a
// NaN cell -- adding +0 to a sum of non-negative terms leaves it
// bit-identical, so the chain is one dependent add per stock. LIN / LOUT are
// the layouts (-1: read from the view at run time).
//
// Many rows (batch runs): the sums go to scratch (4 rows per workgroup, the
// chain being the bound) and a second kernel divides every cell on all CUs;
// one row (stream ticks, pool segments): one workgroup sums and divides.
constexpr int SCALE_BLOCK = 256;
constexpr int SCALE_ROWS = 8;
constexpr int SCALE_SUM_ROWS = 4;
constexpr int SCALE_TILE = 1024;

template <int LAY> __device__ __forceinline__ size_t lay_at(const kun_cs::View &v, long row, long s, long S, int blk) {
    if (LAY == 1) return (size_t)row * S + s;
    if (LAY == 2) return (size_t)s;
    return cs_at(v.layout, row, s, S, v.rows, blk);
}

// sums != null: store the row sums there and stop
template <typename T, int LIN, int LOUT, int ROWS>
__device__ __forceinline__ void scale_block(const kun_cs::View &in, const kun_cs::View &out, long S, long rows,
                                            int blk, const long r0, T *sums) {
    constexpr int PER = ROWS * SCALE_TILE / SCALE_BLOCK;
    __shared__ T tile[ROWS * (SCALE_TILE + 1)];
    __shared__ T total[ROWS];
    const int nr = (int)((rows - r0) < ROWS ? (rows - r0) : ROWS);
    const T *src = (const T *)in.ptr;
    T *dst = (T *)out.ptr;
    T reg[PER];
    auto load = [&](long c0) {
#pragma unroll
        for (int q = 0; q < PER; q++) {
            const int e = threadIdx.x + q * SCALE_BLOCK;
            const int r = e / SCALE_TILE, c = e % SCALE_TILE;
            const long s = c0 + c;
            reg[q] = (r < nr && s < S) ? src[lay_at<LIN>(in, r0 + r + in.row0, s, S, blk)] : T(0);
        }
    };
    load(0);
    T sum = T(0);
    for (long c0 = 0; c0 < S; c0 += SCALE_TILE) {
        __syncthreads();
#pragma unroll
        for (int q = 0; q < PER; q++) {
            const int e = threadIdx.x + q * SCALE_BLOCK;
            const T x = reg[q];
            tile[(e / SCALE_TILE) * (SCALE_TILE + 1) + e % SCALE_TILE] = (x == x) ? kun::vabs(x) : T(0);
        }
        __syncthreads();
        if (c0 + SCALE_TILE < S) load(c0 + SCALE_TILE);
        if ((int)threadIdx.x < nr) {
            const int n = (int)((S - c0) < SCALE_TILE ? (S - c0) : SCALE_TILE);
            const T *row = tile + threadIdx.x * (SCALE_TILE + 1);
            // in stock order, as the reference's scalar loop; each group's
            // LDS reads are issued before its adds, so their latency
            // overlaps the chain
            int k = 0;
            for (; k + 16 <= n; k += 16) {
                T g[16];
#pragma unroll
                for (int u = 0; u < 16; u++) g[u] = row[k + u];
#pragma unroll
                for (int u = 0; u < 16; u++) sum = sum + g[u];
            }
            for (; k < n; k++) sum = sum + row[k];
        }
    }
    if (sums) {
        if ((int)threadIdx.x < nr) sums[r0 + threadIdx.x] = sum;
        return;
    }
    if ((int)threadIdx.x < nr) total[threadIdx.x] = sum;
    __syncthreads();
    for (int r = 0; r < nr; r++) {
        const T tot = total[r];
        const long rr = r0 + r;
        for (long s = threadIdx.x; s < S; s += SCALE_BLOCK) {
            const T x = src[lay_at<LIN>(in, rr + in.row0, s, S, blk)];
            dst[lay_at<LOUT>(out, rr + out.row0, s, S, blk)] = (x == T(0) && tot == T(0)) ? kun::qnan<T>() : x / tot;
        }
    }
}

template <typename T, int LIN, int LOUT>
__global__ __launch_bounds__(SCALE_BLOCK) void scale_rows(kun_cs::View in, kun_cs::View out, long S, long rows,
                                                           int blk) {
    scale_block<T, LIN, LOUT, SCALE_ROWS>(in, out, S, rows, blk, (long)blockIdx.x * SCALE_ROWS, nullptr);
}

template <typename T, int LIN>
__global__ __launch_bounds__(SCALE_BLOCK) void scale_sums(kun_cs::View in, long S, long rows, int blk, T *sums) {
    scale_block<T, LIN, LIN, SCALE_SUM_ROWS>(in, in, S, rows, blk, (long)blockIdx.x * SCALE_SUM_ROWS, sums);
}

// out = x / sum of its row, or NaN where x == 0 and the sum is 0; one cell
// per thread over rows x S
template <typename T, int LIN, int LOUT>
__global__ __launch_bounds__(256) void scale_div(kun_cs::View in, kun_cs::View out, long S, long rows, int blk,
                                                 const T *sums) {
    const long i = (long)blockIdx.x * 256 + threadIdx.x;
    if (i >= rows * S) return;
    const long r = i / S, s = i - r * S;
    const T x = ((const T *)in.ptr)[lay_at<LIN>(in, r + in.row0, s, S, blk)];
    const T tot = sums[r];
    ((T *)out.ptr)[lay_at<LOUT>(out, r + out.row0, s, S, blk)] = (x == T(0) && tot == T(0)) ? kun::qnan<T>() : x / tot;
}

template <typename T> __global__ __launch_bounds__(SCALE_BLOCK) void scale_segs(const kun_cs::Seg *segs) {
    const kun_cs::Seg g = segs[blockIdx.x];
    const kun_cs::View in{const_cast<void *>(g.in), 2, 1, 0}, out{g.out, 2, 1, 0};
    scale_block<T, 2, 2, SCALE_ROWS>(in, out, g.n, 1, 1, 0, nullptr);
}

template <typename T, int LIN, int LOUT>
hipError_t launch_scale_split(const kun_cs::View &in, const kun_cs::View &out, long S, long rows, int blk,
                              hipStream_t st, T *sums) {
    hipLaunchKernelGGL((scale_sums<T, LIN>), dim3((unsigned)((rows + SCALE_SUM_ROWS - 1) / SCALE_SUM_ROWS)),
                       dim3(SCALE_BLOCK), 0, st, in, S, rows, blk, sums);
    hipLaunchKernelGGL((scale_div<T, LIN, LOUT>), dim3((unsigned)((rows * S + 255) / 256)), dim3(256), 0, st, in, out,
                       S, rows, blk, (const T *)sums);
    return hipGetLastError();
}

template <typename T>
hipError_t launch_scale_t(const kun_cs::View &in, const kun_cs::View &out, long S, long rows, int blk,
                          hipStream_t st, void *scratch) {
    if (scratch && rows > 1) {
        T *sums = (T *)scratch;
        if (in.layout == 1 && out.layout == 1) return launch_scale_split<T, 1, 1>(in, out, S, rows, blk, st, sums);
        return launch_scale_split<T, -1, -1>(in, out, S, rows, blk, st, sums);
    }
    const unsigned grid = (unsigned)((rows + SCALE_ROWS - 1) / SCALE_ROWS);
    if (in.layout == 1 && out.layout == 1)
        hipLaunchKernelGGL((scale_rows<T, 1, 1>), dim3(grid), dim3(SCALE_BLOCK), 0, st, in, out, S, rows, blk);
    else if (in.layout == 2 && out.layout == 2)
        hipLaunchKernelGGL((scale_rows<T, 2, 2>), dim3(grid), dim3(SCALE_BLOCK), 0, st, in, out, S, rows, blk);
    else
        hipLaunchKernelGGL((scale_rows<T, -1, -1>), dim3(grid), dim3(SCALE_BLOCK), 0, st, in, out, S, rows, blk);
    return hipGetLastError();
}

template <typename T, int BLOCK, int ITEMS, bool MERGE>
hipError_t launch_rank_n(const kun_cs::View &in, const kun_cs::View &out, long S, long rows, int blk,
                         hipStream_t st) {
    hipLaunchKernelGGL((rank_rows<T, BLOCK, ITEMS, MERGE>), dim3((unsigned)rows), dim3(BLOCK), 0, st, in, out, S,
                       blk);
    return hipGetLastError();
}

template <typename T>
hipError_t launch_rank_t(const kun_cs::View &in, const kun_cs::View &out, long S, long rows, int blk,
                         hipStream_t st) {
    if (S <= 256) return launch_rank_n<T, 256, 1, false>(in, out, S, rows, blk, st);
    if (S <= 512) return launch_rank_n<T, 256, 2, false>(in, out, S, rows, blk, st);
    if (S <= 1024) return launch_rank_n<T, 256, 4, false>(in, out, S, rows, blk, st);
    if (S <= 2048) return launch_rank_n<T, 1024, 2, true>(in, out, S, rows, blk, st);
    if (S <= 4096) return launch_rank_n<T, 1024, 4, true>(in, out, S, rows, blk, st);
    if (S <= 8192) {
        // 64-bit keys: moving no stock payload through the merge pays for the
        // binary searches (fp32: the pair form is faster)
        if constexpr (sizeof(T) == 8) {
            hipLaunchKernelGGL((rank_rows_keys<T, 1024, 8>), dim3((unsigned)rows), dim3(1024), 0, st, in, out, S, blk);
            return hipGetLastError();
        } else {
            return launch_rank_n<T, 1024, 8, true>(in, out, S, rows, blk, st);
        }
    }
    if (S <= 16384) {
        // a 64-bit-key merge sort of 16384 pairs needs more LDS than a CU has
        if constexpr (sizeof(T) == 8)
            return launch_rank_n<T, 512, 32, false>(in, out, S, rows, blk, st);
        else
            return launch_rank_n<T, 1024, 16, true>(in, out, S, rows, blk, st);
    }
    return hipErrorInvalidValue;
}

template <typename T, int BLOCK, int ITEMS, bool MERGE>
hipError_t launch_rank_segs_n(const kun_cs::Seg *segs, int nseg, hipStream_t st) {
    hipLaunchKernelGGL((rank_segs<T, BLOCK, ITEMS, MERGE>), dim3((unsigned)nseg), dim3(BLOCK), 0, st, segs);
    return hipGetLastError();
}

template <typename T> hipError_t launch_rank_segs_t(const kun_cs::Seg *segs, int nseg, long maxS, hipStream_t st) {
    if (maxS <= 256) return launch_rank_segs_n<T, 256, 1, false>(segs, nseg, st);
    if (maxS <= 512) return launch_rank_segs_n<T, 256, 2, false>(segs, nseg, st);
    if (maxS <= 1024) return launch_rank_segs_n<T, 256, 4, false>(segs, nseg, st);
    if (maxS <= 2048) return launch_rank_segs_n<T, 1024, 2, true>(segs, nseg, st);
    if (maxS <= 4096) return launch_rank_segs_n<T, 1024, 4, true>(segs, nseg, st);
    if (maxS <= 8192) return launch_rank_segs_n<T, 1024, 8, true>(segs, nseg, st);
    if (maxS <= 16384) {
        if constexpr (sizeof(T) == 8)
            return launch_rank_segs_n<T, 512, 32, false>(segs, nseg, st);
        else
            return launch_rank_segs_n<T, 1024, 16, true>(segs, nseg, st);
    }
    return hipErrorInvalidValue;
}

} // namespace

namespace kun_cs {

hipError_t launch_rank_segs(int dtype, const Seg *segs, int nseg, long maxS, hipStream_t st) {
    if (nseg <= 0) return hipSuccess;
    return dtype ? launch_rank_segs_t<double>(segs, nseg, maxS, st) : launch_rank_segs_t<float>(segs, nseg, maxS, st);
}

hipError_t launch_scale_segs(int dtype, const Seg *segs, int nseg, hipStream_t st) {
    if (nseg <= 0) return hipSuccess;
    if (dtype)
        hipLaunchKernelGGL((scale_segs<double>), dim3((unsigned)nseg), dim3(SCALE_BLOCK), 0, st, segs);
    else
        hipLaunchKernelGGL((scale_segs<float>), dim3((unsigned)nseg), dim3(SCALE_BLOCK), 0, st, segs);
    return hipGetLastError();
}


long max_rank_stocks(int dtype) { return 16384; }

hipError_t launch_rank(int dtype, const View &in, const View &out, long S, long rows, int blk, hipStream_t st) {
    if (rows <= 0) return hipSuccess;
    return dtype ? launch_rank_t<double>(in, out, S, rows, blk, st) : launch_rank_t<float>(in, out, S, rows, blk, st);
}

hipError_t launch_scale(int dtype, const View &in, const View &out, long S, long rows, int blk, hipStream_t st,
                        void *scratch) {
    if (rows <= 0) return hipSuccess;
    return dtype ? launch_scale_t<double>(in, out, S, rows, blk, st, scratch)
                 : launch_scale_t<float>(in, out, S, rows, blk, st, scratch);
}

} // namespace kun_cs
