// Cross-sectional row kernels: Rank and Scale over the stocks of one row.
//
// Rank (restates cpp/Kun/Rank.hpp:11-57): per row, the non-NaN values are
// sorted; a value's rank is the average 1-based position of its tie group
// [lo, hi): sum = (lo + hi + 1) * (hi - lo) / 2 in integers, then
// T(sum) / T(hi - lo) / T(n). -0 and +0 tie (std::sort / equal_range use
// operator<). On the device a row's values become order-preserving integer
// keys (-0 canonicalised to +0, NaN to the largest key) and are sorted with
// their stocks by an LSD radix sort in LDS, 4-bit digits counted per lane
// (CountRank, below); tie groups are read off the sorted order.
//
// Scale (restates cpp/Kun/Scale.hpp:13-46): sum of |x| over non-NaN values
// accumulated sequentially in stock order (bit-exact order), then
// x / sum, or NaN when x == 0 and sum == 0.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <tuple>
#include <type_traits>
#include <utility>

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

// LAY: the view's layout fixed at compile time (-1: read from the view)
template <int LAY> __device__ __forceinline__ size_t lay_at(const kun_cs::View &v, long row, long s, long S, int blk) {
    if (LAY == 1) return (size_t)row * S + s;
    if (LAY == 2) return (size_t)s;
    return cs_at(v.layout, row, s, S, v.rows, blk);
}

// A kernel launch through a sink: the arguments converted to the kernel's
// parameter types, one pointer to each.
template <typename... P, size_t... I>
hipError_t sink_launch_impl(kun_cs::Sink &sink, const void *k, dim3 g, dim3 b, std::index_sequence<I...>,
                            std::tuple<P...> vals) {
    void *ptrs[] = {(void *)&std::get<I>(vals)...};
    return sink.launch(k, g, b, ptrs);
}
template <typename... P, typename... A>
hipError_t sink_launch(kun_cs::Sink &sink, void (*k)(P...), dim3 g, dim3 b, A... a) {
    static_assert(sizeof...(P) == sizeof...(A), "kernel argument count");
    return sink_launch_impl<P...>(sink, (const void *)k, g, b, std::index_sequence_for<P...>{},
                                  std::tuple<P...>((P)a...));
}

// ---- wavefront primitives (64 lanes) ------------------------------------

// number of lanes below this one whose bit is set in `mask`
__device__ __forceinline__ uint32_t lanes_below(uint64_t mask) {
    return __builtin_amdgcn_mbcnt_hi((uint32_t)(mask >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)mask, 0u));
}

// lanes of the wavefront holding the same 8-bit digit as this lane: one
// ballot per digit bit; a lane disagrees with the lanes whose ballot bit
// differs from its own bit (m: all ones when this lane's bit is set), and
// the peers are the lanes that never disagree
// (bits on which the row's keys differ, `vary`, uniform: the others tell no
// lanes apart and are skipped)
__device__ __forceinline__ uint64_t same_digit(uint32_t d, uint32_t vary) {
    uint32_t miss_lo = 0, miss_hi = 0;
#pragma unroll
    for (int b = 0; b < 8; ++b) {
        if (!((vary >> b) & 1u)) continue;
        const uint32_t m = (uint32_t)((int32_t)(d << (31 - b)) >> 31);
        const uint64_t bal = __ballot(m != 0u);
        miss_lo |= (uint32_t)bal ^ m;
        miss_hi |= (uint32_t)(bal >> 32) ^ m;
    }
    return ~(((uint64_t)miss_hi << 32) | miss_lo);
}

template <typename V> __device__ __forceinline__ V wave_incl_sum(V v) {
    const int lane = threadIdx.x & 63;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        const V u = __shfl_up(v, o, 64);
        if (lane >= o) v += u;
    }
    return v;
}

template <typename V> __device__ __forceinline__ V wave_incl_max(V v) {
    const int lane = threadIdx.x & 63;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        const V u = __shfl_up(v, o, 64);
        if (lane >= o && u > v) v = u;
    }
    return v;
}

template <typename V> __device__ __forceinline__ V wave_sum(V v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}
template <typename V> __device__ __forceinline__ V wave_and(V v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v &= __shfl_xor(v, o, 64);
    return v;
}
template <typename V> __device__ __forceinline__ V wave_or(V v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v |= __shfl_xor(v, o, 64);
    return v;
}

// rank of every member of the tie group at sorted positions [a, b) among n
// non-NaN values
template <typename T> __device__ __forceinline__ T group_rank(int a, int b, int n) {
    // rows of at most 16384 stocks: (a + b + 1) * (b - a) < 2^31, so the
    // reference's integer sum fits 32 bits and converts to T the same way
    const int32_t sum = (a + b + 1) * (b - a) / 2;
    // x / 1 == x exactly: a group of one skips the first division
    const T q = (b - a == 1) ? (T)sum : (T)sum / (T)(b - a);
    return q / (T)n;
}

// ---- Rank of long 64-bit rows: keys-only LSD radix sort -------------------
//
// For 64-bit keys of rows above 8192 stocks, whose (key, stock) pairs would
// not fit in LDS. A workgroup of BLOCK lanes (NW = BLOCK / 64 wavefronts)
// holds N = BLOCK x ITEMS keys in registers, wavefront w owning positions
// [w*64*ITEMS, (w+1)*64*ITEMS): item j of lane l is position
// w*64*ITEMS + j*64 + l (loads and stores of consecutive lanes are
// consecutive). Each pass sorts stably by one 8-bit digit:
//  1. ranking: per item, the lanes with the same digit are found by ballots
//     (same_digit); the lowest such lane adds their count to the wavefront's
//     own LDS histogram row, every lane's rank within its digit being the
//     row's previous count plus its lanes_below;
//  2. a block scan turns the histograms into starting positions, digit by
//     digit and within a digit wavefront by wavefront -- the order the
//     positions held -- so the pass is stable;
//  3. keys are scattered to LDS and read back.
// Digit positions on which every key of the row agrees are skipped (their
// stable pass would keep the order); the row's AND and OR of keys tell.
// Each lane then finds its own keys' tie groups by binary search in the
// sorted keys.
template <typename T, int BLOCK, int ITEMS> struct KeysRank {
    using K = typename KeyOf<T>::K;
    static constexpr int N = BLOCK * ITEMS;
    static constexpr int NW = BLOCK / 64;
    static_assert(NW % 4 == 0 && NW <= 16, "BLOCK must be 256, 512 or 1024");
    struct Smem {
        K key[N];                        // keys in sorted order
        uint32_t hist[NW][256];          // per wavefront digit counts, then starting positions
        uint32_t wv[NW];                 // per wavefront totals of the scans
        K wand[NW], wor[NW];
    };

    // hist := exclusive prefix in (digit, wavefront) order. Thread t owns
    // digit t / (NW/4) of wavefronts 4g .. 4g+3 (g = t % (NW/4)): thread
    // order is that order, so one block scan of per-thread sums does it.
    static __device__ __forceinline__ void scan_hist(Smem &sm) {
        constexpr int G = NW / 4;
        const int t = threadIdx.x, d = t / G, g = t % G, lane = t & 63, w = t >> 6;
        uint32_t v[4], s = 0;
#pragma unroll
        for (int i = 0; i < 4; i++) {
            v[i] = sm.hist[4 * g + i][d];
            s += v[i];
        }
        const uint32_t incl = wave_incl_sum(s);
        if (lane == 63) sm.wv[w] = incl;
        __syncthreads();
        uint32_t run = incl - s;
        for (int q = 0; q < w; q++) run += sm.wv[q];
#pragma unroll
        for (int i = 0; i < 4; i++) {
            sm.hist[4 * g + i][d] = run;
            run += v[i];
        }
    }

    static __device__ __forceinline__ void row(Smem &sm, const kun_cs::View &in, const kun_cs::View &out, long S,
                                               int blk, long t) {
        if (S <= 0) return; // uniform: nothing to rank
        const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
        const int base = w * 64 * ITEMS + lane;
        const T *src = (const T *)in.ptr;
        T *dst = (T *)out.ptr;
        const long rin = t + in.row0;
        K key[ITEMS];
        int cnt = 0;
        K kand = ~K(0), kor = K(0);
        // every load issued before any is used (stocks past S read the last
        // stock's cell and are then replaced by padding keys)
        T xs[ITEMS];
#pragma unroll
        for (int j = 0; j < ITEMS; j++) {
            const long s = base + j * 64;
            xs[j] = src[cs_at(in.layout, rin, s < S ? s : S - 1, S, in.rows, blk)];
        }
#pragma unroll
        for (int j = 0; j < ITEMS; j++) {
            const long s = base + j * 64;
            const bool live = s < S;
            const K k = live ? rank_key(xs[j]) : ~K(0);
            cnt += (live && xs[j] == xs[j]) ? 1 : 0;
            key[j] = k;
            kand &= k;
            kor |= k;
        }
        cnt = wave_sum(cnt);
        kand = wave_and(kand);
        kor = wave_or(kor);
        if (lane == 0) {
            sm.wv[w] = (uint32_t)cnt;
            sm.wand[w] = kand;
            sm.wor[w] = kor;
        }
        __syncthreads();
        int n = 0;
#pragma unroll
        for (int q = 0; q < NW; q++) {
            n += (int)sm.wv[q];
            kand &= sm.wand[q];
            kor |= sm.wor[q];
        }
        const K diff = kand ^ kor; // bits on which the row's keys differ
        __syncthreads();           // sm.wv is the scans' from here on
        bool moved = false;
        for (int p = 0; p < (int)sizeof(K); p++) {
            const int sh = 8 * p;
            if (((diff >> sh) & K(0xff)) == K(0)) continue; // uniform over the workgroup
            moved = true;
            uint32_t *hw = sm.hist[w];
            // this wavefront's row only: its reads of the row in the
            // previous pass ended before the last barrier
            for (int q = lane; q < 256; q += 64) hw[q] = 0;
            // rank of each key within its digit among this wavefront's keys
            // before it, two 16-bit ranks per register
            uint32_t loc2[(ITEMS + 1) / 2];
            const uint32_t vary = (uint32_t)(diff >> sh) & 0xffu;
#pragma unroll
            for (int j = 0; j < ITEMS; j++) {
                const uint32_t d = (uint32_t)(key[j] >> sh) & 0xffu;
                const uint64_t peers = same_digit(d, vary);
                const uint32_t below = lanes_below(peers);
                const uint32_t prior = hw[d];
                if (below == 0) hw[d] = prior + (uint32_t)__popcll(peers);
                if (j % 2 == 0) loc2[j / 2] = prior + below;
                else loc2[j / 2] |= (prior + below) << 16;
                // one key at a time: interleaving the keys' ballot chains
                // buys no speed and costs the registers of a second
                // resident workgroup
                __builtin_amdgcn_sched_barrier(0);
            }
            __syncthreads();
            scan_hist(sm);
            __syncthreads();
#pragma unroll
            for (int j = 0; j < ITEMS; j++) {
                const uint32_t pos = hw[(uint32_t)(key[j] >> sh) & 0xffu] + ((loc2[j / 2] >> (16 * (j % 2))) & 0xffffu);
                sm.key[pos] = key[j];
            }
            __syncthreads();
#pragma unroll
            for (int j = 0; j < ITEMS; j++) key[j] = sm.key[base + j * 64];
        }
        if (!moved) {
            // every key equal: the order is sorted as it stands
#pragma unroll
            for (int j = 0; j < ITEMS; j++) sm.key[base + j * 64] = key[j];
        }
        __syncthreads(); // the sorted keys are in sm.key
        // this lane's own keys again (the registers now hold sorted ones)
#pragma unroll
        for (int j = 0; j < ITEMS; j++) {
            const long s = base + j * 64;
            key[j] = rank_key(src[cs_at(in.layout, rin, s < S ? s : S - 1, S, in.rows, blk)]);
        }
#pragma unroll
        for (int j = 0; j < ITEMS; j++)
            if (base + j * 64 >= S) key[j] = ~K(0);
        rank_by_search(sm, key, n, dst, out, S, blk, t, base);
    }

    static __device__ __forceinline__ T rank_of(int a, int b, int n) { return group_rank<T>(a, b, n); }

    static __device__ __forceinline__ void rank_by_search(Smem &sm, const K (&mine)[ITEMS], int n, T *dst,
                                                          const kun_cs::View &out, long S, int blk, long t,
                                                          int base) {
        // groups of G keys at a time, their searches interleaved
        constexpr int G = ITEMS < 4 ? ITEMS : 4;
#pragma unroll
        for (int j0 = 0; j0 < ITEMS; j0 += G) {
            // lower bounds
            int lo[G], len[G];
#pragma unroll
            for (int j = 0; j < G; j++) {
                lo[j] = 0;
                len[j] = n;
            }
            for (int step = 0; step < 32; step++) {
                bool any = false;
#pragma unroll
                for (int j = 0; j < G; j++) {
                    if (len[j] > 0) {
                        const int half = len[j] >> 1;
                        if (sm.key[lo[j] + half] < mine[j0 + j]) {
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
            // upper bounds, searched only for keys with ties
            int hi[G];
#pragma unroll
            for (int j = 0; j < G; j++) {
                hi[j] = lo[j] + 1;
                len[j] = (mine[j0 + j] != ~K(0) && hi[j] < n && sm.key[hi[j]] == mine[j0 + j]) ? n - hi[j] : 0;
            }
            for (int step = 0; step < 32; step++) {
                bool any = false;
#pragma unroll
                for (int j = 0; j < G; j++) {
                    if (len[j] > 0) {
                        const int half = len[j] >> 1;
                        if (!(mine[j0 + j] < sm.key[hi[j] + half])) {
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
            for (int j = 0; j < G; j++) {
                const long s = base + (j0 + j) * 64;
                if (s >= S) continue;
                const T r = (mine[j0 + j] != ~K(0) && lo[j] < n) ? rank_of(lo[j], hi[j], n) : kun::qnan<T>();
                dst[cs_at(out.layout, t + out.row0, s, S, out.rows, blk)] = r;
            }
        }
    }
};

// inclusive prefix sum / max over the 64 lanes by DPP row shifts and row
// broadcasts (rows of 16 lanes, then across rows); max assumes values >= 0
__device__ __forceinline__ uint32_t dpp_incl_sum(uint32_t v) {
    v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x111, 0xf, 0xf, false); // row_shr:1
    v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x112, 0xf, 0xf, false); // row_shr:2
    v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x114, 0xf, 0xf, false); // row_shr:4
    v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x118, 0xf, 0xf, false); // row_shr:8
    v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x142, 0xa, 0xf, false); // row_bcast:15
    v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x143, 0xc, 0xf, false); // row_bcast:31
    return v;
}
// inclusive prefix sum within each row of 16 lanes
__device__ __forceinline__ uint32_t dpp_row_incl_sum(uint32_t v) {
    v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x111, 0xf, 0xf, false); // row_shr:1
    v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x112, 0xf, 0xf, false); // row_shr:2
    v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x114, 0xf, 0xf, false); // row_shr:4
    v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x118, 0xf, 0xf, false); // row_shr:8
    return v;
}
__device__ __forceinline__ int dpp_incl_max(int v) {
    v = max(v, __builtin_amdgcn_update_dpp(0, v, 0x111, 0xf, 0xf, false));
    v = max(v, __builtin_amdgcn_update_dpp(0, v, 0x112, 0xf, 0xf, false));
    v = max(v, __builtin_amdgcn_update_dpp(0, v, 0x114, 0xf, 0xf, false));
    v = max(v, __builtin_amdgcn_update_dpp(0, v, 0x118, 0xf, 0xf, false));
    v = max(v, __builtin_amdgcn_update_dpp(0, v, 0x142, 0xa, 0xf, false));
    v = max(v, __builtin_amdgcn_update_dpp(0, v, 0x143, 0xc, 0xf, false));
    return v;
}

// ---- Rank: LSD radix sort with per-lane digit counters -------------------
//
// The (key, stock) form for rows whose pairs fit in LDS. A workgroup of
// BLOCK lanes holds N = BLOCK x ITEMS keys in registers, lane t owning
// positions [t*ITEMS, (t+1)*ITEMS) (blocked; the first pass may start from
// any order, the loads are striped for coalescing). Each pass sorts stably by
// one 4-bit digit:
//  1. each lane counts its own keys' digits in registers, sixteen 4-bit
//     fields per 8 keys: a key's rank among the lane's earlier keys of its
//     digit is the fields' value before its increment. The counts go to the
//     lane's own column of the LDS table cnt[8][BLOCK], digit d in half
//     d >> 3 of word cnt[d & 7][t] -- no atomics, no ballots;
//  2. one block scan of the table in memory order (word-major, lane-minor:
//     digits 0-7 in the low halves, 8-15 in the high halves; adding the low
//     digits' total to every high half puts digits 8-15 after them) gives
//     each (digit, lane) the position of its first key;
//  3. keys and stocks are scattered to their positions in LDS and read back
//     blocked.
// Within a digit, positions go lane by lane and within a lane key by key --
// the order the positions held -- so each pass is stable. Digit positions
// on which every key of the row agrees are skipped (the row's AND and OR of
// keys tell). Tie groups come from a max-scan of group starts over the
// sorted positions.
template <typename T, int BLOCK, int ITEMS> struct CountRank {
    using K = typename KeyOf<T>::K;
    static constexpr int N = BLOCK * ITEMS;
    static constexpr int NW = BLOCK / 64;
    static constexpr int NG = (ITEMS + 7) / 8; // groups of up to 8 keys, one field register each
    static constexpr int NB = (int)sizeof(K) * 8;
    static_assert(BLOCK % 64 == 0 && NW <= 16, "BLOCK: 64..1024 lanes");
    static_assert(N <= 16384, "positions and counts are 16-bit");
    struct alignas(16) Smem {
        K key[N];                // sorted keys; scratch of the block scans; then group ends, ranks by stock
        uint16_t idx[N];         // stock at each position
        uint32_t cnt[8 * BLOCK]; // digit counts, then first positions; scratch of the group scan
    };

    static __device__ __forceinline__ uint32_t digit(K k, int sh) { return (uint32_t)(k >> sh) & 15u; }

    template <int LIN, int LOUT>
    static __device__ __forceinline__ void row(Smem &sm, const kun_cs::View &in, const kun_cs::View &out, long S,
                                               int blk, long t) {
        if (S <= 0) return; // uniform: nothing to rank
        const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
        const int wu = __builtin_amdgcn_readfirstlane(w); // w, known uniform
        const T *src = (const T *)in.ptr;
        T *dst = (T *)out.ptr;
        const long rin = t + in.row0;
        T xs[ITEMS];
#pragma unroll
        for (int j = 0; j < ITEMS; j++) {
            const long s = w * 64 * ITEMS + j * 64 + lane;
            xs[j] = src[lay_at<LIN>(in, rin, s < S ? s : S - 1, S, blk)];
        }
        K key[ITEMS];
        uint32_t id2[(ITEMS + 1) / 2]; // stock of each key, two 16-bit ids per register
        int cnt = 0;
        K kand = ~K(0), kor = K(0);
#pragma unroll
        for (int j = 0; j < ITEMS; j++) {
            const long s = w * 64 * ITEMS + j * 64 + lane;
            const bool live = s < S;
            const K k = live ? rank_key(xs[j]) : ~K(0);
            cnt += (live && xs[j] == xs[j]) ? 1 : 0;
            key[j] = k;
            if (j % 2 == 0) id2[j / 2] = (uint32_t)s;
            else id2[j / 2] |= (uint32_t)s << 16;
            kand &= k;
            kor |= k;
        }
        cnt = wave_sum(cnt);
        kand = wave_and(kand);
        kor = wave_or(kor);
        // wavefront partials in the key area (its next writer, the first
        // pass's scan, comes after a barrier)
        K *xa = sm.key, *xo = sm.key + NW;
        uint32_t *xc = (uint32_t *)(sm.key + 2 * NW);
        if (lane == 0) {
            xa[w] = kand;
            xo[w] = kor;
            xc[w] = (uint32_t)cnt;
        }
        __syncthreads();
        // lane q < NW holds wavefront q's partials
        const int n = wave_sum(lane < NW ? (int)xc[lane] : 0);
        kand = wave_and(lane < NW ? xa[lane] : ~K(0));
        kor = wave_or(lane < NW ? xo[lane] : K(0));
        const K diff = kand ^ kor; // bits on which the row's keys differ
        const uint16_t *first16 = (const uint16_t *)sm.cnt;
        uint32_t *xs32 = (uint32_t *)sm.key; // scan scratch
        bool moved = false;
        for (int sh = 0; sh < NB; sh += 4) {
            if (digit(diff, sh) == 0u) continue; // uniform over the workgroup
            moved = true;
            // 1. per-lane counts; pre[j]: the earlier keys of this lane with
            // key j's digit
            uint64_t f[NG];
            uint32_t pre[ITEMS];
#pragma unroll
            for (int g = 0; g < NG; g++) f[g] = 0;
#pragma unroll
            for (int j = 0; j < ITEMS; j++) {
                const uint32_t d4 = digit(key[j], sh) * 4u;
                uint32_t p = 0;
#pragma unroll
                for (int g = 0; g < j / 8; g++) p += (uint32_t)(f[g] >> d4) & 15u;
                pre[j] = p + ((uint32_t)(f[j / 8] >> d4) & 15u);
                f[j / 8] += (uint64_t)1 << d4;
            }
            // kept as computed here rather than recomputed at the scatter
            // from eight 64-bit partial counts
#pragma unroll
            for (int j = 0; j < ITEMS; j++) asm volatile("" : "+v"(pre[j]));
#pragma unroll
            for (int k = 0; k < 8; k++) {
                uint32_t word = 0;
#pragma unroll
                for (int g = 0; g < NG; g++)
                    word += (((uint32_t)f[g] >> (4 * k)) & 15u) | ((((uint32_t)(f[g] >> 32) >> (4 * k)) & 15u) << 16);
                sm.cnt[k * BLOCK + tid] = word;
            }
            __syncthreads();
            // 2. exclusive scan of the table in memory order; lane t owns words [8t, 8t+8)
            uint4 *c4 = (uint4 *)sm.cnt;
            const uint4 lo = c4[2 * tid], hi = c4[2 * tid + 1];
            const uint32_t v[8] = {lo.x, lo.y, lo.z, lo.w, hi.x, hi.y, hi.z, hi.w};
            uint32_t s = 0;
#pragma unroll
            for (int i = 0; i < 8; i++) s += v[i];
            const uint32_t incl = dpp_incl_sum(s);
            if (lane == 63) xs32[w] = incl;
            __syncthreads();
            // wavefront totals: lane q of every wavefront holds the sum of
            // totals 0..q (prefix over the first 16 lanes)
            const uint32_t wp = dpp_row_incl_sum(lane < NW ? xs32[lane] : 0u);
            const uint32_t total = (uint32_t)__builtin_amdgcn_readlane((int)wp, NW - 1);
            const uint32_t before = wu ? (uint32_t)__builtin_amdgcn_readlane((int)wp, wu - 1) : 0u;
            uint32_t run = before + incl - s + ((total & 0xffffu) << 16);
            uint32_t o[8];
#pragma unroll
            for (int i = 0; i < 8; i++) {
                o[i] = run;
                run += v[i];
            }
            c4[2 * tid] = make_uint4(o[0], o[1], o[2], o[3]);
            c4[2 * tid + 1] = make_uint4(o[4], o[5], o[6], o[7]);
            __syncthreads();
            // 3. scatter to the sorted positions (every lookup issued before
            // the first store), read back blocked
            uint32_t pos[ITEMS];
#pragma unroll
            for (int j = 0; j < ITEMS; j++) {
                const uint32_t d = digit(key[j], sh);
                pos[j] = first16[((d & 7u) * BLOCK + tid) * 2 + (d >> 3)];
            }
#pragma unroll
            for (int j = 0; j < ITEMS; j++) {
                const uint32_t at = pos[j] + pre[j];
                sm.key[at] = key[j];
                sm.idx[at] = (uint16_t)(id2[j / 2] >> (16 * (j % 2)));
            }
            __syncthreads();
#pragma unroll
            for (int j = 0; j < ITEMS; j++) key[j] = sm.key[tid * ITEMS + j];
            if constexpr (ITEMS % 2 == 0) {
                // stock pairs (2j, 2j+1) are adjacent: one word each
#pragma unroll
                for (int j = 0; j < ITEMS / 2; j++) id2[j] = ((const uint32_t *)sm.idx)[tid * ITEMS / 2 + j];
            } else {
                id2[0] = sm.idx[tid];
            }
        }
        if (!moved) {
            // every key equal: any order is sorted
            __syncthreads(); // the partials' readers are done
#pragma unroll
            for (int j = 0; j < ITEMS; j++) sm.key[tid * ITEMS + j] = key[j];
            __syncthreads();
        }
        ties<LOUT>(sm, key, id2, n, dst, out, S, blk, t);
    }

    static __device__ __forceinline__ T rank_of(int a, int b, int n) { return group_rank<T>(a, b, n); }

    // sorted keys at positions P = tid*ITEMS + j (registers and sm.key)
    template <int LOUT>
    static __device__ __forceinline__ void ties(Smem &sm, const K (&key)[ITEMS], const uint32_t (&id2)[(ITEMS + 1) / 2],
                                                int n, T *dst, const kun_cs::View &out, long S, int blk, long t) {
        const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
        const int P0 = tid * ITEMS;
        // a[j]: start of key j's group (the last position <= P0 + j that starts one)
        int a[ITEMS];
        bool last[ITEMS];
        int cur = -1;
#pragma unroll
        for (int j = 0; j < ITEMS; j++) {
            const bool first = j > 0 ? key[j - 1] != key[j] : (P0 == 0 || sm.key[P0 - 1] != key[0]);
            last[j] = j + 1 < ITEMS ? key[j] != key[j + 1] : (P0 + ITEMS == N || sm.key[P0 + ITEMS] != key[j]);
            if (first) cur = P0 + j;
            a[j] = cur;
        }
        // group start carried in from earlier lanes: exclusive max-scan of
        // (cur + 1), 0 meaning none
        const int incl = dpp_incl_max(cur + 1);
        int prev = __shfl_up(incl, 1, 64);
        if (lane == 0) prev = 0;
        int *xm = (int *)sm.cnt;
        if (lane == 63) xm[w] = incl;
        __syncthreads(); // also: every read of sm.key above is done
        int carry = prev;
#pragma unroll
        for (int q = 0; q < NW; q++) carry = q < w && xm[q] > carry ? xm[q] : carry;
        carry -= 1;
        int *endof = (int *)sm.key;
#pragma unroll
        for (int j = 0; j < ITEMS; j++) {
            if (a[j] < 0) a[j] = carry;
            const int P = P0 + j;
            if (P < n && last[j]) endof[a[j]] = P + 1;
        }
        __syncthreads();
        T r[ITEMS];
#pragma unroll
        for (int j = 0; j < ITEMS; j++) r[j] = P0 + j < n ? rank_of(a[j], endof[a[j]], n) : kun::qnan<T>();
        __syncthreads();
        T *res = (T *)sm.key;
#pragma unroll
        for (int j = 0; j < ITEMS; j++) {
            const int id = (int)((id2[j / 2] >> (16 * (j % 2))) & 0xffffu);
            if (id < S) res[id] = r[j];
        }
        __syncthreads();
        for (long s = tid; s < S; s += BLOCK) dst[lay_at<LOUT>(out, t + out.row0, s, S, blk)] = res[s];
    }
};

// KIND 0: CountRank (pairs); 1: KeysRank (64-bit keys beyond 8192 stocks)
template <typename T, int BLOCK, int ITEMS, int KIND>
__device__ __forceinline__ void rank_row_k(const kun_cs::View &in, const kun_cs::View &out, long S, int blk,
                                           const long t) {
    if constexpr (KIND == 0) {
        using R = CountRank<T, BLOCK, ITEMS>;
        __shared__ typename R::Smem sm;
        // rows of TS buffers (every batch run's temporaries) with their
        // addressing fixed at compile time; other layouts read the view's
        if (in.layout == 1 && out.layout == 1) R::template row<1, 1>(sm, in, out, S, blk, t);
        else R::template row<-1, -1>(sm, in, out, S, blk, t);
    } else {
        using R = KeysRank<T, BLOCK, ITEMS>;
        __shared__ typename R::Smem sm;
        R::row(sm, in, out, S, blk, t);
    }
}

// one workgroup per row of the views
template <typename T, int BLOCK, int ITEMS, int KIND>
__global__ __launch_bounds__(BLOCK) void rank_rows(kun_cs::View in, kun_cs::View out, long S, int blk) {
    rank_row_k<T, BLOCK, ITEMS, KIND>(in, out, S, blk, blockIdx.x);
}

// one workgroup per segment (a contiguous run of stocks of one row)
template <typename T, int BLOCK, int ITEMS, int KIND>
__global__ __launch_bounds__(BLOCK) void rank_segs(const kun_cs::Seg *segs) {
    const kun_cs::Seg g = segs[blockIdx.x];
    const kun_cs::View in{const_cast<void *>(g.in), 2, 1, 0}, out{g.out, 2, 1, 0};
    rank_row_k<T, BLOCK, ITEMS, KIND>(in, out, g.n, 1, 0);
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

// Scratch of the split form: the rows' sums, then (float rows) RN(1/sum) in
// double at the next multiple of 8 bytes
template <typename T> __device__ __forceinline__ double *scale_rcs(T *sums, long rows) {
    return (double *)((char *)sums + ((size_t)rows * sizeof(T) + 7) / 8 * 8);
}

// x / tot, exactly as IEEE division, or NaN where x == 0 and tot == 0; float
// rows multiply by rc = RN(1/tot) in double (kun::div_rc)
__device__ __forceinline__ float scale_q(float x, float tot, double rc) {
    const float q = kun::div_rc(x, rc, tot);
    return (x == 0.f && tot == 0.f) ? kun::qnan<float>() : q;
}
__device__ __forceinline__ double scale_q(double x, double tot, double) {
    return (x == 0.0 && tot == 0.0) ? kun::qnan<double>() : x / tot;
}

// sums != null: store the row sums there (and their reciprocals, float)
// and stop
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
        if ((int)threadIdx.x < nr) {
            sums[r0 + threadIdx.x] = sum;
            if (sizeof(T) == 4) scale_rcs(sums, rows)[r0 + threadIdx.x] = 1.0 / (double)sum;
        }
        return;
    }
    if ((int)threadIdx.x < nr) total[threadIdx.x] = sum;
    __syncthreads();
    for (int r = 0; r < nr; r++) {
        const T tot = total[r];
        const double rc = sizeof(T) == 4 ? 1.0 / (double)tot : 0.0;
        const long rr = r0 + r;
        for (long s = threadIdx.x; s < S; s += SCALE_BLOCK) {
            const T x = src[lay_at<LIN>(in, rr + in.row0, s, S, blk)];
            dst[lay_at<LOUT>(out, rr + out.row0, s, S, blk)] = scale_q(x, tot, rc);
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

// out = x / sum of its row, or NaN where x == 0 and the sum is 0: workgroup
// (x, y) takes SCALE_DIV_ITEMS x 256 stocks of rows y, y + gridDim.y, ...
constexpr int SCALE_DIV_ITEMS = 4;
template <typename T, int LIN, int LOUT>
__global__ __launch_bounds__(256) void scale_div(kun_cs::View in, kun_cs::View out, long S, long rows, int blk,
                                                 T *sums) {
    const long s0 = (long)blockIdx.x * (256 * SCALE_DIV_ITEMS) + threadIdx.x;
    const T *src = (const T *)in.ptr;
    T *dst = (T *)out.ptr;
    auto rows_of = [&](auto full) {
        for (long r = blockIdx.y; r < rows; r += gridDim.y) {
            const T tot = sums[r];
            const double rc = sizeof(T) == 4 ? scale_rcs(sums, rows)[r] : 0.0;
            T x[SCALE_DIV_ITEMS];
#pragma unroll
            for (int q = 0; q < SCALE_DIV_ITEMS; q++) {
                const long s = s0 + q * 256;
                x[q] = (full || s < S) ? src[lay_at<LIN>(in, r + in.row0, s, S, blk)] : T(0);
            }
#pragma unroll
            for (int q = 0; q < SCALE_DIV_ITEMS; q++) {
                const long s = s0 + q * 256;
                if (full || s < S) dst[lay_at<LOUT>(out, r + out.row0, s, S, blk)] = scale_q(x[q], tot, rc);
            }
        }
    };
    // whole workgroups of stocks take the unguarded copy
    if ((long)(blockIdx.x + 1) * (256 * SCALE_DIV_ITEMS) <= S)
        rows_of(std::true_type{});
    else
        rows_of(std::false_type{});
}

template <typename T> __global__ __launch_bounds__(SCALE_BLOCK) void scale_segs(const kun_cs::Seg *segs) {
    const kun_cs::Seg g = segs[blockIdx.x];
    const kun_cs::View in{const_cast<void *>(g.in), 2, 1, 0}, out{g.out, 2, 1, 0};
    scale_block<T, 2, 2, SCALE_ROWS>(in, out, g.n, 1, 1, 0, nullptr);
}

template <typename T, int LIN, int LOUT>
hipError_t launch_scale_split(const kun_cs::View &in, const kun_cs::View &out, long S, long rows, int blk,
                              kun_cs::Sink &sink, T *sums) {
    hipError_t e = sink_launch(sink, scale_sums<T, LIN>, dim3((unsigned)((rows + SCALE_SUM_ROWS - 1) / SCALE_SUM_ROWS)),
                               dim3(SCALE_BLOCK), in, S, rows, blk, sums);
    if (e != hipSuccess) return e;
    const long per = 256L * SCALE_DIV_ITEMS;
    return sink_launch(sink, scale_div<T, LIN, LOUT>,
                       dim3((unsigned)((S + per - 1) / per), (unsigned)(rows < 65535 ? rows : 65535)), dim3(256), in,
                       out, S, rows, blk, sums);
}

template <typename T>
hipError_t launch_scale_t(const kun_cs::View &in, const kun_cs::View &out, long S, long rows, int blk,
                          kun_cs::Sink &sink, void *scratch) {
    if (scratch && rows > 1) {
        T *sums = (T *)scratch;
        if (in.layout == 1 && out.layout == 1) return launch_scale_split<T, 1, 1>(in, out, S, rows, blk, sink, sums);
        return launch_scale_split<T, -1, -1>(in, out, S, rows, blk, sink, sums);
    }
    const dim3 grid((unsigned)((rows + SCALE_ROWS - 1) / SCALE_ROWS));
    if (in.layout == 1 && out.layout == 1)
        return sink_launch(sink, scale_rows<T, 1, 1>, grid, dim3(SCALE_BLOCK), in, out, S, rows, blk);
    if (in.layout == 2 && out.layout == 2)
        return sink_launch(sink, scale_rows<T, 2, 2>, grid, dim3(SCALE_BLOCK), in, out, S, rows, blk);
    return sink_launch(sink, scale_rows<T, -1, -1>, grid, dim3(SCALE_BLOCK), in, out, S, rows, blk);
}

template <typename T, int BLOCK, int ITEMS, int KIND>
hipError_t launch_rank_n(const kun_cs::View &in, const kun_cs::View &out, long S, long rows, int blk,
                         kun_cs::Sink &sink) {
    return sink_launch(sink, rank_rows<T, BLOCK, ITEMS, KIND>, dim3((unsigned)rows), dim3(BLOCK), in, out, S, blk);
}

// Workgroup shape by row length: at most 1024 lanes, up to 16 keys per lane;
// 64-bit keys of more than 8192 stocks sort without payloads (their pairs
// would not fit in LDS).
template <typename T>
hipError_t launch_rank_t(const kun_cs::View &in, const kun_cs::View &out, long S, long rows, int blk,
                         kun_cs::Sink &sink) {
    if (S <= 256) return launch_rank_n<T, 256, 1, 0>(in, out, S, rows, blk, sink);
    if (S <= 512) return launch_rank_n<T, 256, 2, 0>(in, out, S, rows, blk, sink);
    if (S <= 1024) return launch_rank_n<T, 256, 4, 0>(in, out, S, rows, blk, sink);
    if (S <= 2048) return launch_rank_n<T, 512, 4, 0>(in, out, S, rows, blk, sink);
    if (S <= 4096) return launch_rank_n<T, 1024, 4, 0>(in, out, S, rows, blk, sink);
    if (S <= 8192) return launch_rank_n<T, 1024, 8, 0>(in, out, S, rows, blk, sink);
    if (S <= 16384) return launch_rank_n<T, 1024, 16, sizeof(T) == 8 ? 1 : 0>(in, out, S, rows, blk, sink);
    return hipErrorInvalidValue;
}

template <typename T, int BLOCK, int ITEMS, int KIND>
hipError_t launch_rank_segs_n(const kun_cs::Seg *segs, int nseg, hipStream_t st) {
    hipLaunchKernelGGL((rank_segs<T, BLOCK, ITEMS, KIND>), dim3((unsigned)nseg), dim3(BLOCK), 0, st, segs);
    return hipGetLastError();
}

template <typename T> hipError_t launch_rank_segs_t(const kun_cs::Seg *segs, int nseg, long maxS, hipStream_t st) {
    if (maxS <= 256) return launch_rank_segs_n<T, 256, 1, 0>(segs, nseg, st);
    if (maxS <= 512) return launch_rank_segs_n<T, 256, 2, 0>(segs, nseg, st);
    if (maxS <= 1024) return launch_rank_segs_n<T, 256, 4, 0>(segs, nseg, st);
    if (maxS <= 2048) return launch_rank_segs_n<T, 512, 4, 0>(segs, nseg, st);
    if (maxS <= 4096) return launch_rank_segs_n<T, 1024, 4, 0>(segs, nseg, st);
    if (maxS <= 8192) return launch_rank_segs_n<T, 1024, 8, 0>(segs, nseg, st);
    if (maxS <= 16384) return launch_rank_segs_n<T, 1024, 16, sizeof(T) == 8 ? 1 : 0>(segs, nseg, st);
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

hipError_t launch_rank(int dtype, const View &in, const View &out, long S, long rows, int blk, Sink &sink) {
    if (rows <= 0) return hipSuccess;
    return dtype ? launch_rank_t<double>(in, out, S, rows, blk, sink)
                 : launch_rank_t<float>(in, out, S, rows, blk, sink);
}

hipError_t launch_rank(int dtype, const View &in, const View &out, long S, long rows, int blk, hipStream_t st) {
    StreamSink sink(st);
    return launch_rank(dtype, in, out, S, rows, blk, sink);
}

size_t scale_scratch_bytes(int dtype, long rows) {
    const size_t n = (size_t)(rows > 1 ? rows : 1);
    return dtype ? n * 8 : (n * 4 + 7) / 8 * 8 + n * 8;
}

hipError_t launch_scale(int dtype, const View &in, const View &out, long S, long rows, int blk, Sink &sink,
                        void *scratch) {
    if (rows <= 0) return hipSuccess;
    return dtype ? launch_scale_t<double>(in, out, S, rows, blk, sink, scratch)
                 : launch_scale_t<float>(in, out, S, rows, blk, sink, scratch);
}

hipError_t launch_scale(int dtype, const View &in, const View &out, long S, long rows, int blk, hipStream_t st,
                        void *scratch) {
    StreamSink sink(st);
    return launch_scale(dtype, in, out, S, rows, blk, sink, scratch);
}

} // namespace kun_cs
