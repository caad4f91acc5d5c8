// Device templates for generated factor kernels (gfx950 / CDNA4).
//
// One lane owns one stock; a time step of a TS buffer is one coalesced row.
// Every primitive reproduces the lane semantics of the reference CPU runtime
// bit for bit (cited per function). Generated code must be compiled with
// -ffp-contract=off: the only fused multiply-adds are the explicit kun::fma
// calls that mirror the reference's sc_fmadd.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "kun_logtables.h"

namespace kun {

template <typename T> struct Bits;
template <> struct Bits<float> {
    using U = uint32_t;
    using I = int32_t;
    static constexpr int MANT = 23, EXPB = 8, NBITS = 32;
};
template <> struct Bits<double> {
    using U = uint64_t;
    using I = int64_t;
    static constexpr int MANT = 52, EXPB = 11, NBITS = 64;
};

template <typename T> __device__ __forceinline__ typename Bits<T>::U to_bits(T x) {
    return __builtin_bit_cast(typename Bits<T>::U, x);
}
template <typename T> __device__ __forceinline__ T from_bits(typename Bits<T>::U u) {
    return __builtin_bit_cast(T, u);
}
template <typename T> __device__ __forceinline__ T qnan();
template <> __device__ __forceinline__ float qnan<float>() { return __builtin_nanf(""); }
template <> __device__ __forceinline__ double qnan<double>() { return __builtin_nan(""); }
template <typename T> __device__ __forceinline__ T inf() { return (T)__builtin_huge_val(); }

// comparison results are all-ones / all-zero bit masks (x86 _CMP_*_OQ)
template <typename T> __device__ __forceinline__ T mask(bool c) {
    using U = typename Bits<T>::U;
    return from_bits<T>(c ? ~U(0) : U(0));
}
template <typename T> __device__ __forceinline__ bool sign_set(T x) {
    return (to_bits(x) >> (Bits<T>::NBITS - 1)) & 1;
}
// blendv: picks a where the sign bit of c is set (f32x8.hpp:84-87)
template <typename T> __device__ __forceinline__ T sel(T c, T a, T b) { return sign_set(c) ? a : b; }
template <typename T> __device__ __forceinline__ T bor(T a, T b) { return from_bits<T>(to_bits(a) | to_bits(b)); }
template <typename T> __device__ __forceinline__ T band(T a, T b) { return from_bits<T>(to_bits(a) & to_bits(b)); }
template <typename T> __device__ __forceinline__ T bnot(T a) { return from_bits<T>(~to_bits(a)); }
// _mm256_min_ps / _mm256_max_ps: second operand unless the first compares
// strictly (NaN or equal operands give b) (Ops.hpp:618-626)
template <typename T> __device__ __forceinline__ T vmin(T a, T b) { return a < b ? a : b; }
template <typename T> __device__ __forceinline__ T vmax(T a, T b) { return a > b ? a : b; }
template <typename T> __device__ __forceinline__ T vabs(T a) {
    using U = typename Bits<T>::U;
    return from_bits<T>(to_bits(a) & ~(U(1) << (Bits<T>::NBITS - 1)));
}
// x / c for a float c, rounded exactly as IEEE float division, given
// rc = RN(1/c) in double. The double product x * rc is within 2^-52 of x / c
// (relative). A normal float result is then exact: x / c is never a
// midpoint between normal floats (a midpoint's odd part has 25 bits, x's at
// most 24) and lies at least 2^-49 (relative) from one, so rounding the
// product to float rounds x / c. A subnormal result is not: midpoints there
// have short odd parts, x / c can be exactly one (c = 16350, x = 2047 * 8175
// * 2^-149), and the product's error then picks a side instead of the even
// neighbour. Those results (and only those: a zero product is never the
// wrong side of a tie) take the division. Checked for every float x and a
// set of divisors, tie-prone ones included, by
// tests/native/div_const_check.c.
__device__ __forceinline__ float div_rc(float x, double rc, float c) {
    const float r = (float)((double)x * rc);
    if (__builtin_expect(__builtin_isfpclass(r, 0x0090), 0)) return x / c;  // subnormal
    return r;
}
// the bare product, for divisors whose midpoint quotients all round right
// (every float x checked when the kernel is generated)
__device__ __forceinline__ float div_const(float x, double rc, float) { return (float)((double)x * rc); }
__device__ __forceinline__ float vsqrt(float a) { return __builtin_sqrtf(a); }
__device__ __forceinline__ double vsqrt(double a) { return __builtin_sqrt(a); }
__device__ __forceinline__ float fma(float a, float b, float c) { return __builtin_fmaf(a, b, c); }
__device__ __forceinline__ double fma(double a, double b, double c) { return __builtin_fma(a, b, c); }
__device__ __forceinline__ float vfloor(float a) { return __builtin_floorf(a); }
__device__ __forceinline__ double vfloor(double a) { return __builtin_floor(a); }

// Sign (Ops.hpp:692-701)
template <typename T> __device__ __forceinline__ T vsign(T v) {
    T r = (v != v) ? qnan<T>() : T(1);
    r = v < T(0) ? T(-1) : r;
    r = v == T(0) ? T(0) : r;
    return r;
}
// SetInfOrNanToValue (Ops.hpp:715-720)
template <typename T> __device__ __forceinline__ T set_inf_nan(T a, T v) {
    T d = a - a;
    return d != d ? v : a;
}

// LogFast: table + degree-5 polynomial (KunSIMD/cpu/Math.hpp:20-93)
template <typename T> struct LogTables;
template <> struct LogTables<float> {
    static __device__ __forceinline__ float r(unsigned i) {
        constexpr float t[32] = KUN_LOG_R_F32;
        return t[i];
    }
    static __device__ __forceinline__ float lr(unsigned i) {
        constexpr float t[32] = KUN_LOG_LR_F32;
        return t[i];
    }
};
template <> struct LogTables<double> {
    static __device__ __forceinline__ double r(unsigned i) {
        constexpr double t[32] = KUN_LOG_R_F64;
        return t[i];
    }
    static __device__ __forceinline__ double lr(unsigned i) {
        constexpr double t[32] = KUN_LOG_LR_F64;
        return t[i];
    }
};

template <typename T> __device__ __forceinline__ T log_fast(T x) {
    using B = Bits<T>;
    using U = typename B::U;
    constexpr int approx = 5;
    U xi = to_bits(x);
    U hi5 = (xi >> (B::MANT - approx)) & U(31);
    U hi1 = hi5 >> (approx - 1);
    U e1 = (hi1 ^ ((U(1) << (B::EXPB - 1)) - 1)) << B::MANT;
    U m = (xi & ((U(1) << B::MANT) - 1)) | e1;
    T ri = LogTables<T>::r((unsigned)hi5);
    T z = ri * from_bits<T>(m);
    z = z - T(1.0);
    T p = T(1 / 5.0);
    p = fma(p, z, T(-1 / 4.0));
    p = fma(p, z, T(1 / 3.0));
    p = fma(p, z, T(-1 / 2.0));
    p = fma(p, z, T(1.0));
    p = p * z;
    T lr = LogTables<T>::lr((unsigned)hi5);
    typename B::I er = (typename B::I)((xi >> B::MANT) + hi1);
    T v2 = fma((T)er, T(0.6931471805599453), lr);
    T hi = p + v2;
    T lo = hi - v2;
    lo = lo - p;
    hi = hi + lo;
    hi = x == T(0) ? -inf<T>() : hi;
    hi = x < T(0) ? qnan<T>() : hi;
    hi = x == inf<T>() ? inf<T>() : hi;
    hi = x != x ? qnan<T>() : hi;
    hi = x == T(1) ? T(0) : hi;
    return hi;
}

// natural log of the C library, used when fast_log is disabled
__device__ __forceinline__ float log_lib(float x) { return __builtin_logf(x); }
__device__ __forceinline__ double log_lib(double x) { return __builtin_log(x); }

// Exp: range reduction + degree-6 Horner (KunSIMD/cpu/Math.hpp:96-139)
template <typename T> __device__ __forceinline__ T exp_fast(T a) {
    using B = Bits<T>;
    using U = typename B::U;
    using I = typename B::I;
    T kf = vfloor(fma(a, T(1.4426950408889634), T(0.5)));
    const T lim = T(Bits<T>::NBITS == 32 ? 2.0e9 : 9.0e18);
    I ki = (kf != kf || kf >= lim || kf <= -lim) ? I(0) : (I)kf;
    T r = fma(-kf, T(0.6931471805599453), a);
    T t = fma(r, T(0.16666666666666666), T(1.0));
    t = fma(t, r * T(0.2), T(1.0));
    t = fma(t, r * T(0.25), T(1.0));
    t = fma(t, r * T(0.3333333333333333), T(1.0));
    t = fma(t, r * T(0.5), T(1.0));
    t = fma(t, r, T(1.0));
    U res = ((U)ki << B::MANT) + to_bits(t);
    T out = from_bits<T>(res);
    out = a > T(88.72283935) ? inf<T>() : out;
    out = a < T(-87.33654785) ? T(0) : out;
    return out;
}

// Kahan-compensated rolling sum (FastWindowedSum, Ops.hpp:324-356)
template <typename T> struct FastSum {
    T v, cadd, csub;
    typename Bits<T>::I nnan;
    __device__ __forceinline__ void init(int window) {
        v = T(0);
        cadd = T(0);
        csub = T(0);
        nnan = window;
    }
    __device__ __forceinline__ T step(T old, T cur) {
        bool onan = old != old, cnan = cur != cur;
        if (!onan) {
            T y = T(0) - old;
            y = y - csub;
            T t = v + y;
            csub = (t - v) - y;
            v = t;
        }
        if (!cnan) {
            T y = cur - cadd;
            T t = v + y;
            cadd = (t - v) - y;
            v = t;
        }
        nnan = nnan - (onan ? 1 : 0) + (cnan ? 1 : 0);
        return nnan == 0 ? v : qnan<T>();
    }
};

// ewm(span=w, adjust=False, ignore_na=True) (Ops.hpp:358-375)
template <typename T> struct Ema {
    T v;
    __device__ __forceinline__ T step(T cur, T w, T keep) {
        T nv = v * keep;
        T add = cur * w;
        nv = nv + add;
        T r = cur != cur ? v : nv;
        r = v != v ? cur : r;
        v = r;
        return r;
    }
};

// rolling regression sums against x = 1..w (Ops.hpp:377-466)
template <typename T> struct LinReg {
    T is, xs, x2, ys, y2, xy;
    typename Bits<T>::I nnan;
    __device__ __forceinline__ void init(int window) {
        is = xs = x2 = ys = y2 = xy = T(0);
        nnan = window;
    }
    __device__ __forceinline__ void step(T old, T cur, T W) {
        bool onan = old != old, cnan = cur != cur;
        xy = xy - ys;
        x2 = (x2 + is) - T(2.0) * xs;
        xs = xs - is;
        if (!onan) {
            is = is - T(1);
            ys = ys - old;
            y2 = y2 - old * old;
        }
        nnan = nnan - (onan ? 1 : 0) + (cnan ? 1 : 0);
        if (!cnan) {
            is = is + T(1);
            xs = xs + W;
            x2 = x2 + W * W;
            ys = ys + cur;
            y2 = y2 + cur * cur;
            xy = xy + W * cur;
        }
    }
    __device__ __forceinline__ T rsquare(int window) const {
        T N = (T)(typename Bits<T>::I)(window - nnan);
        T r1 = N * xy - xs * ys;
        r1 = r1 * r1;
        T r2 = (N * x2 - xs * xs) * (N * y2 - ys * ys);
        return r1 / r2;
    }
    __device__ __forceinline__ T slope(int window) const {
        T N = (T)(typename Bits<T>::I)(window - nnan);
        return (N * xy - xs * ys) / (N * x2 - xs * xs);
    }
    __device__ __forceinline__ T resi(int window, T val) const {
        T N = (T)(typename Bits<T>::I)(window - nnan);
        T sl = slope(window);
        T xm = xs / N, ym = ys / N;
        T interp = ym - sl * xm;
        return val - (sl * (T)window + interp);
    }
};

// Sorted window of the last W non-NaN values of one lane, stored
// column-wise ([W+1][stride], LDS in batch kernels, HBM in stream kernels);
// restates the skip-list state of Ops/Quantile.hpp:15-45. Insertion goes
// after equal values and removal takes the first equal value, so ties keep
// insertion order like the reference skip list.
template <typename T, int W> struct SortedWin {
    T *val;      // [W + 1][stride]
    int32_t *ix; // [W + 1][stride]
    long stride;
    int n, last;
    __device__ __forceinline__ void init(T *v, int32_t *i, long stride_) {
        val = v;
        ix = i;
        stride = stride_;
        n = 0;
        last = 0;
    }
    __device__ __forceinline__ T get(int k, long lane) const { return val[k * stride + lane]; }
    __device__ __forceinline__ int32_t idx(int k, long lane) const { return ix[k * stride + lane]; }
    __device__ __forceinline__ void step(T old, T cur, int32_t index, long lane) {
        if (old == old) {
            int p = 0;
            while (p < n && get(p, lane) < old) p++;
            if (p < n && get(p, lane) == old) {
                for (int q = p; q < n - 1; q++) {
                    val[q * stride + lane] = val[(q + 1) * stride + lane];
                    ix[q * stride + lane] = ix[(q + 1) * stride + lane];
                }
                n--;
            }
        }
        if (cur == cur) {
            int p = 0;
            while (p < n && get(p, lane) <= cur) p++;
            for (int q = n; q > p; q--) {
                val[q * stride + lane] = val[(q - 1) * stride + lane];
                ix[q * stride + lane] = ix[(q - 1) * stride + lane];
            }
            val[p * stride + lane] = cur;
            ix[p * stride + lane] = index;
            n++;
            last = p;
        }
    }
    // SkipListQuantile (Ops/Quantile.hpp:84-105): q*(n-1) in T, interpolation in double
    __device__ __forceinline__ T quantile(T q, long lane) const {
        if (n != W) return qnan<T>();
        T f = q * (T)(n - 1);
        int i = (int)f;
        if ((T)i == f) return get(i, lane);
        double lo = (double)get(i, lane), hi = (double)get(i + 1, lane);
        return (T)(lo + (hi - lo) * (double)(f - (T)i));
    }
    // SkipListRank (Ops/Quantile.hpp:107-128)
    __device__ __forceinline__ T rank(T cur, long lane) const {
        if (n != W) return qnan<T>();
        int lt = 0;
        for (int k = 0; k < n; k++) lt += get(k, lane) < cur ? 1 : 0;
        double r = last + 1, rmin = lt + 1;
        r = (((r * (r + 1) / 2) - ((rmin - 1) * rmin / 2)) / (r - rmin + 1));
        return (T)r;
    }
    __device__ __forceinline__ T vmin_(long lane) const { return n != W ? qnan<T>() : get(0, lane); }
    __device__ __forceinline__ T vmax_(long lane) const { return n != W ? qnan<T>() : get(W - 1, lane); }
    __device__ __forceinline__ T argmin(int64_t cur_idx, long lane) const {
        if (n != W) return qnan<T>();
        return (T)(int64_t)(idx(0, lane) + n - cur_idx);
    }
};

// The same sorted window held in registers by a lane that walks every row
// of the run (sequential batch kernels): every access uses a compile-time
// slot, so insertion and removal are W + 1 selects instead of a chain of
// dependent memory round trips. IDX keeps the row indices (SkipListArgMin).
// Ordering is that of SortedWin: a new value goes after equal ones, a
// removal takes the first equal one.
template <typename T, int W, bool IDX> struct SortedWinR {
    T v[W + 1];
    int32_t ix[IDX ? W + 1 : 1];
    int n = 0, last = 0;
    __device__ __forceinline__ SortedWinR() {
#pragma unroll
        for (int k = 0; k <= W; k++) v[k] = T(0);
#pragma unroll
        for (int k = 0; k < (IDX ? W + 1 : 1); k++) ix[k] = 0;
    }
    __device__ __forceinline__ void step(T old, T cur, int32_t index, long) {
        if (old == old) {
            int p = 0;
#pragma unroll
            for (int k = 0; k <= W; k++) p += (k < n && v[k] < old) ? 1 : 0;
            bool found = false;
#pragma unroll
            for (int k = 0; k <= W; k++) found = found || (k == p && k < n && v[k] == old);
            if (found) {
#pragma unroll
                for (int k = 0; k < W; k++) {
                    v[k] = k >= p ? v[k + 1] : v[k];
                    if (IDX) ix[k] = k >= p ? ix[k + 1] : ix[k];
                }
                n--;
            }
        }
        if (cur == cur) {
            int p = 0;
#pragma unroll
            for (int k = 0; k <= W; k++) p += (k < n && v[k] <= cur) ? 1 : 0;
#pragma unroll
            for (int k = W; k > 0; k--) {
                v[k] = k > p ? v[k - 1] : v[k];
                if (IDX) ix[k] = k > p ? ix[k - 1] : ix[k];
            }
#pragma unroll
            for (int k = 0; k <= W; k++) {
                v[k] = k == p ? cur : v[k];
                if (IDX) ix[k] = k == p ? index : ix[k];
            }
            n++;
            last = p;
        }
    }
    __device__ __forceinline__ T at(int i) const {
        T r = T(0);
#pragma unroll
        for (int k = 0; k <= W; k++) r = k == i ? v[k] : r;
        return r;
    }
    __device__ __forceinline__ T quantile(T q, long) const {
        if (n != W) return qnan<T>();
        T f = q * (T)(n - 1);
        int i = (int)f;
        if ((T)i == f) return at(i);
        double lo = (double)at(i), hi = (double)at(i + 1);
        return (T)(lo + (hi - lo) * (double)(f - (T)i));
    }
    __device__ __forceinline__ T rank(T cur, long) const {
        if (n != W) return qnan<T>();
        int lt = 0;
#pragma unroll
        for (int k = 0; k < W; k++) lt += v[k] < cur ? 1 : 0;
        double r = last + 1, rmin = lt + 1;
        r = (((r * (r + 1) / 2) - ((rmin - 1) * rmin / 2)) / (r - rmin + 1));
        return (T)r;
    }
    __device__ __forceinline__ T vmin_(long) const { return n != W ? qnan<T>() : v[0]; }
    __device__ __forceinline__ T vmax_(long) const { return n != W ? qnan<T>() : v[W - 1]; }
    __device__ __forceinline__ T argmin(int64_t cur_idx, long) const {
        if (n != W) return qnan<T>();
        return (T)(int64_t)(ix[0] + n - cur_idx);
    }
};

// `v` loaded from row `row` (clamped at 0 by the caller) is NaN when the row
// lies before the run start -- no history before start (Ops.hpp:87-92)
template <typename T> __device__ __forceinline__ T nan_before(T v, long row) { return row >= 0 ? v : qnan<T>(); }

// Window read of a [rows][S] buffer at a row known to lie inside it: a raw
// buffer load whose byte offset is the wave-uniform `roff` (row * S *
// sizeof(T), kept in a scalar register) plus the lane's `so` (stock *
// sizeof(T)), so the address costs no vector instruction. The caller
// guarantees roff + so < 2^31. Word 3 of the descriptor is the gfx9
// raw-buffer format; num_records is the maximum, so no load is clipped.
template <typename T> __device__ __forceinline__ T wld(const T *base, unsigned so, long roff) {
    const __amdgpu_buffer_rsrc_t r = __builtin_amdgcn_make_buffer_rsrc((void *)base, (short)0, -1, 0x00020000);
    if constexpr (sizeof(T) == 4) {
        return __builtin_bit_cast(T, __builtin_amdgcn_raw_buffer_load_b32(r, (int)so, (int)roff, 0));
    } else {
        return __builtin_bit_cast(T, __builtin_amdgcn_raw_buffer_load_b64(r, (int)so, (int)roff, 0));
    }
}

// element (row, stock) of a buffer in TS [rows][S] or STs [S/b][rows][b]
template <int LAYOUT> __device__ __forceinline__ size_t at(long row, long s, long S, long rows, int blk) {
    if (LAYOUT == 1) return (size_t)row * S + s;
    return ((size_t)(s / blk) * rows + row) * blk + s % blk;
}

#ifndef KUN_MAX_STAGE_BUFS
#define KUN_MAX_STAGE_BUFS 256
#endif

// Arguments of a batch stage kernel, passed by value: the stage's buffers in
// the order of its descriptor (inputs, then outputs), then its materialised
// values.
#ifndef KUN_MAX_COMPS
#define KUN_MAX_COMPS 64
#endif

struct StageArgs {
    void *bufs[KUN_MAX_STAGE_BUFS];
    void *scratch;     // HBM columns of rings that do not fit in LDS, sorted windows
    long S;            // stocks
    long T;            // rows of the user inputs
    long start;        // first row of user inputs to compute
    long L;            // rows to compute
    int blk;           // STs blocking length
    // launch table: entry e covers workgroups [zfirst[e], zfirst[e+1]) and
    // runs component zcomp[e] over znch[e] chunks of zchunk[e] rows, each
    // chunk x 64-stock block one workgroup; its scratch starts at zscr[e]
    int nz;
    int zcomp[KUN_MAX_COMPS];
    int zfirst[KUN_MAX_COMPS + 1];
    int zchunk[KUN_MAX_COMPS];
    int znch[KUN_MAX_COMPS];
    long zscr[KUN_MAX_COMPS];
};

} // namespace kun

namespace kun {
// Arguments of a one-tick stream stage kernel: current rows of the stage's
// buffers, the stage's carried state (history rings, op state) as
// structure-of-arrays [slot][S] in HBM, and the tick counter.
struct StreamArgs {
    void *bufs[KUN_MAX_STAGE_BUFS];
    void *st;     // T slots
    int32_t *ist; // integer slots
    long S;
    long tick;
};
// Generic cross-sectional stage (a GenericCrossSectionalOp's body compiled
// as device code): one thread per row. Buffer n (stage inputs, then
// outputs) is addressed in layout lay[n] -- 0 STs (rows[n] rows, blocking
// blk), 1 TS, 2 a single row -- and computed row r is its row row0[n] + r.
#define KUN_MAX_CS_BUFS 16
struct CsArgs {
    void *bufs[KUN_MAX_CS_BUFS];
    int lay[KUN_MAX_CS_BUFS];
    long rows[KUN_MAX_CS_BUFS];
    long row0[KUN_MAX_CS_BUFS];
    long S;
    long L;
    int blk;
};
// input_<n> / output_<n> of a body: element i = stock i of one row
template <typename T> struct CsRow {
    T *p;
    int lay;
    long S, rows, row;
    int blk;
    __device__ __forceinline__ T &operator[](size_t i) const {
        const long s = (long)i;
        const size_t at = lay == 1   ? (size_t)row * S + s
                          : lay == 2 ? (size_t)s
                                     : ((size_t)(s / blk) * rows + row) * blk + s % blk;
        return p[at];
    }
};
// Session pools: one entry per workgroup of a pool launch -- the lanes
// [s0, s1) of one session's stock range (at most one wavefront), where the
// session's rows start in the packed input rows of the tick (p0), and the
// session's tick counter.
struct PoolBlock {
    int s0, s1;
    long p0;
    long tick;
};
struct PoolArgs {
    const StreamArgs *stages; // per stage of the module (device memory)
    const PoolBlock *blocks;
};
} // namespace kun
