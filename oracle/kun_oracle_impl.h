/* Oracle interpreter body, instantiated once per element type.
 *
 * TEST INFRASTRUCTURE ONLY: this is the CPU restatement used to check the
 * device backend; the product path never links or calls it.
 *
 * Expected macros: REAL (float/double), UBITS (uint32_t/uint64_t),
 * SBITS (int32_t/int64_t), FN(x) name suffixing, MANT (23/52),
 * EXPB (8/11), LOG_R, LOG_LR (table initialisers).
 *
 * Each function cites the reference code whose arithmetic it restates.
 */

static const REAL FN(log_r)[32] = LOG_R;
static const REAL FN(log_lr)[32] = LOG_LR;

static inline UBITS FN(bits)(REAL x) { UBITS u; memcpy(&u, &x, sizeof u); return u; }
static inline REAL FN(from_bits)(UBITS u) { REAL x; memcpy(&x, &u, sizeof x); return x; }
static inline REAL FN(mask)(int c) { return FN(from_bits)(c ? ~(UBITS)0 : (UBITS)0); }
static inline int FN(sign_set)(REAL x) { return (int)(FN(bits)(x) >> (sizeof(UBITS) * 8 - 1)); }

/* LogFast: table + degree-5 polynomial, KunSIMD/cpu/Math.hpp:20-93. */
static REAL FN(log_fast)(REAL x) {
    const int approx = 5;
    UBITS xi = FN(bits)(x);
    UBITS hi5 = (xi >> (MANT - approx)) & 31u;
    UBITS hi1 = hi5 >> (approx - 1);
    UBITS e1 = (hi1 ^ (((UBITS)1 << (EXPB - 1)) - 1)) << MANT;
    UBITS m = (xi & (((UBITS)1 << MANT) - 1)) | e1;
    REAL ri = FN(log_r)[hi5];
    REAL z = ri * FN(from_bits)(m);
    z = z - (REAL)1.0;
    REAL p = (REAL)(1 / 5.0);
    p = FMA(p, z, (REAL)(-1 / 4.0));
    p = FMA(p, z, (REAL)(1 / 3.0));
    p = FMA(p, z, (REAL)(-1 / 2.0));
    p = FMA(p, z, (REAL)1.0);
    p = p * z;
    REAL lr = FN(log_lr)[hi5];
    SBITS er = (SBITS)((xi >> MANT) + hi1);
    REAL ef = (REAL)er;
    REAL v2 = FMA(ef, (REAL)0.6931471805599453, lr);
    REAL hi = p + v2;
    REAL lo = hi - v2;
    lo = lo - p;
    hi = hi + lo;
    if (x == 0) hi = -INFINITY;
    if (x < 0) hi = NAN;
    if (x == INFINITY) hi = INFINITY;
    if (x != x) hi = NAN;
    if (x == 1) hi = 0;
    return hi;
}

/* Exp: range reduction + degree-6 Horner, KunSIMD/cpu/Math.hpp:96-139. */
static REAL FN(exp_fast)(REAL a) {
    REAL kf = FLOOR(FMA(a, (REAL)1.4426950408889634, (REAL)0.5));
    SBITS ki = (kf != kf || kf >= (REAL)9.2e18 || kf <= (REAL)-9.2e18) ? 0 : (SBITS)kf;
    REAL r = FMA(-kf, (REAL)0.6931471805599453, a);
    REAL t = FMA(r, (REAL)0.16666666666666666, (REAL)1.0);
    t = FMA(t, r * (REAL)0.2, (REAL)1.0);
    t = FMA(t, r * (REAL)0.25, (REAL)1.0);
    t = FMA(t, r * (REAL)0.3333333333333333, (REAL)1.0);
    t = FMA(t, r * (REAL)0.5, (REAL)1.0);
    t = FMA(t, r, (REAL)1.0);
    UBITS res = ((UBITS)ki << MANT) + FN(bits)(t);
    REAL out = FN(from_bits)(res);
    if (a > (REAL)88.72283935) out = INFINITY;
    if (a < (REAL)-87.33654785) out = 0;
    return out;
}

static REAL FN(const_of)(const Op *op) {
    return (REAL)PARSE(op->value);
}

/* ------------------------------------------------------------------ */
/* Per-lane-block execution context of one time-series stage.          */

typedef struct {
    REAL *val;     /* [nops][LB] current values */
    REAL **ring;   /* per op: [R][LB] history ring, or NULL */
    int *ringsz;   /* per op: ring rows (power of two) */
    REAL **st;     /* per op: state slots [8][LB] */
    int64_t **cnt; /* per op: integer counter [LB] */
    /* sorted-window state (SkipListState) */
    double **sk_val;  /* [LB][w] */
    int64_t **sk_idx; /* [LB][w] */
    int **sk_n;       /* [LB] */
    int **sk_last;    /* [LB] */
    long rows;        /* rows pushed into the rings so far */
} FN(Ctx);

static void FN(ctx_free)(FN(Ctx) *c, const Stage *sg) {
    if (!c) return;
    for (int k = 0; k < sg->nops; k++) {
        free(c->ring[k]); free(c->st[k]); free(c->cnt[k]);
        free(c->sk_val[k]); free(c->sk_idx[k]); free(c->sk_n[k]); free(c->sk_last[k]);
    }
    free(c->val); free(c->ring); free(c->ringsz); free(c->st); free(c->cnt);
    free(c->sk_val); free(c->sk_idx); free(c->sk_n); free(c->sk_last);
    free(c);
}

static int FN(is_window_source)(const Stage *sg, int k, int *maxoff) {
    int m = -1;
    for (int j = 0; j < sg->nops; j++) {
        const Op *o = &sg->ops[j];
        switch (o->kind) {
        case OP_IterValue:
            if (o->in[1] == k) { int w = sg->ops[o->in[0]].window; if (w - 1 > m) m = w - 1; }
            break;
        case OP_BackRef: case OP_FastWindowedSum: case OP_WindowedLinearRegression:
            if (o->in[0] == k && o->window > m) m = o->window;
            break;
        default: break;
        }
    }
    *maxoff = m;
    return m >= 0;
}

static FN(Ctx) *FN(ctx_new)(const Stage *sg, int LB, const REAL *init_vals /* [nops][LB] or NULL */) {
    FN(Ctx) *c = (FN(Ctx) *)calloc(1, sizeof(FN(Ctx)));
    int n = sg->nops;
    c->val = (REAL *)calloc((size_t)n * LB, sizeof(REAL));
    c->ring = (REAL **)calloc(n, sizeof(REAL *));
    c->ringsz = (int *)calloc(n, sizeof(int));
    c->st = (REAL **)calloc(n, sizeof(REAL *));
    c->cnt = (int64_t **)calloc(n, sizeof(int64_t *));
    c->sk_val = (double **)calloc(n, sizeof(double *));
    c->sk_idx = (int64_t **)calloc(n, sizeof(int64_t *));
    c->sk_n = (int **)calloc(n, sizeof(int *));
    c->sk_last = (int **)calloc(n, sizeof(int *));
    for (int k = 0; k < n; k++) {
        const Op *o = &sg->ops[k];
        int maxoff;
        if (FN(is_window_source)(sg, k, &maxoff)) {
            int r = 1;
            while (r < maxoff + 1) r <<= 1;
            c->ringsz[k] = r;
            c->ring[k] = (REAL *)malloc((size_t)r * LB * sizeof(REAL));
            for (long q = 0; q < (long)r * LB; q++) c->ring[k][q] = NAN;
        }
        switch (o->kind) {
        case OP_FastWindowedSum:
            c->st[k] = (REAL *)calloc((size_t)8 * LB, sizeof(REAL));
            c->cnt[k] = (int64_t *)malloc(LB * sizeof(int64_t));
            for (int l = 0; l < LB; l++) c->cnt[k][l] = o->window;
            break;
        case OP_WindowedLinearRegression:
            c->st[k] = (REAL *)calloc((size_t)8 * LB, sizeof(REAL));
            c->cnt[k] = (int64_t *)malloc(LB * sizeof(int64_t));
            for (int l = 0; l < LB; l++) c->cnt[k][l] = o->window;
            break;
        case OP_ExpMovingAvg:
            c->st[k] = (REAL *)calloc((size_t)8 * LB, sizeof(REAL));
            for (int l = 0; l < LB; l++) c->st[k][l] = NAN;
            break;
        case OP_Accumulator:
            c->st[k] = (REAL *)calloc((size_t)8 * LB, sizeof(REAL));
            break;
        case OP_SkipListState:
            c->sk_val[k] = (double *)calloc((size_t)LB * (o->window + 1), sizeof(double));
            c->sk_idx[k] = (int64_t *)calloc((size_t)LB * (o->window + 1), sizeof(int64_t));
            c->sk_n[k] = (int *)calloc(LB, sizeof(int));
            c->sk_last[k] = (int *)calloc(LB, sizeof(int));
            break;
        default: break;
        }
    }
    (void)init_vals;
    return c;
}

/* ---- buffer access -------------------------------------------------- */
typedef struct {
    void *ptr;
    int layout;     /* 0 STs, 1 TS, 2 row (stream) */
    long num_time;  /* rows stored */
    long base;      /* absolute row of stored row 0 */
} FN(BufView);

static inline size_t FN(buf_index)(const FN(BufView) *b, long row, long s, long S, int blk) {
    if (b->layout == 2) return (size_t)s;
    if (b->layout == 1) return (size_t)row * S + s;
    return ((size_t)(s / blk) * b->num_time + row) * blk + s % blk;
}

/* ---- one time step ---------------------------------------------------- */
typedef struct {
    const Prog *pg;
    const Stage *sg;
    FN(BufView) *views; /* per buffer id */
    long S, start, lane0;
    int LB, nl;         /* block width and active lanes */
} FN(StepEnv);

#define V(k) (c->val + (size_t)(k) * LB)

static inline void FN(ring_push)(FN(Ctx) *c, int k, int LB, const REAL *v, long row) {
    int r = c->ringsz[k];
    REAL *dst = c->ring[k] + (size_t)(row & (r - 1)) * LB;
    memcpy(dst, v, sizeof(REAL) * LB);
}

/* value of window source k, `off` rows before the current row */
static inline const REAL *FN(ring_at)(FN(Ctx) *c, int k, int LB, long row, int off) {
    int r = c->ringsz[k];
    return c->ring[k] + (size_t)((row - off) & (r - 1)) * LB;
}

static void FN(sk_insert)(double *vals, int64_t *idx, int *n, int *last, double v, int64_t index) {
    /* insertion after all equal elements; returns number of elements <= v */
    int lo = 0, hi = *n;
    while (lo < hi) { int mid = (lo + hi) / 2; if (vals[mid] <= v) lo = mid + 1; else hi = mid; }
    for (int q = *n; q > lo; q--) { vals[q] = vals[q - 1]; idx[q] = idx[q - 1]; }
    vals[lo] = v; idx[lo] = index; (*n)++;
    *last = lo;
}

static void FN(sk_remove)(double *vals, int64_t *idx, int *n, double v) {
    /* removes the first element equal to v, if any */
    int lo = 0, hi = *n;
    while (lo < hi) { int mid = (lo + hi) / 2; if (vals[mid] < v) lo = mid + 1; else hi = mid; }
    if (lo >= *n || !(vals[lo] == v)) return;
    for (int q = lo; q < *n - 1; q++) { vals[q] = vals[q + 1]; idx[q] = idx[q + 1]; }
    (*n)--;
}

static void FN(eval_loop)(FN(Ctx) *c, const FN(StepEnv) *e, int L, long row);

static void FN(eval_op)(FN(Ctx) *c, const FN(StepEnv) *e, int k, long row, long i, int iter) {
    const Stage *sg = e->sg;
    const Op *o = &sg->ops[k];
    int LB = e->LB, nl = e->nl;
    REAL *out = V(k);
    const REAL *a = o->nin > 0 ? V(o->in[0]) : NULL;
    const REAL *b = o->nin > 1 ? V(o->in[1]) : NULL;
    const REAL *cc = o->nin > 2 ? V(o->in[2]) : NULL;
    int l;
    switch (o->kind) {
    case OP_Input: {
        const FN(BufView) *bv = &e->views[o->buf];
        const Buf *bi = &e->pg->bufs[o->buf];
        long r;
        if (o->single) r = 0;
        else if (bv->layout == 2) r = 0;
        else r = (bi->kind == 0 ? e->start : 0) + i - bv->base;
        const REAL *p = (const REAL *)bv->ptr;
        for (l = 0; l < nl; l++) {
            long s = e->lane0 + l;
            out[l] = o->single ? p[s] : p[FN(buf_index)(bv, r, s, e->S, e->pg->blocking)];
        }
        for (; l < LB; l++) out[l] = NAN;
        break;
    }
    case OP_Output: {
        FN(BufView) *bv = &e->views[o->buf];
        REAL *p = (REAL *)bv->ptr;
        long r = bv->layout == 2 ? 0 : i;
        for (l = 0; l < nl; l++) p[FN(buf_index)(bv, r, e->lane0 + l, e->S, e->pg->blocking)] = a[l];
        memcpy(out, a, sizeof(REAL) * LB);
        break;
    }
    case OP_WindowedTempOutput: case OP_ReturnFirstValue:
        memcpy(out, a, sizeof(REAL) * LB);
        break;
    case OP_ConstantOp: { REAL v = FN(const_of)(o); for (l = 0; l < LB; l++) out[l] = v; break; }
#define BIN_CONST(EXPR_NOSWAP, EXPR_SWAP) { REAL C_ = FN(const_of)(o); \
        if (!o->swap) { for (l = 0; l < LB; l++) { REAL x = a[l], y = C_; out[l] = EXPR_NOSWAP; } } \
        else { for (l = 0; l < LB; l++) { REAL x = C_, y = a[l]; out[l] = EXPR_SWAP; } } \
        break; }
    case OP_AddConst: BIN_CONST(x + y, x + y)
    case OP_SubConst: BIN_CONST(x - y, x - y)
    case OP_MulConst: BIN_CONST(x * y, x * y)
    case OP_DivConst: BIN_CONST(x / y, x / y)
    case OP_GreaterThanConst: BIN_CONST(FN(mask)(x > y), FN(mask)(x > y))
    case OP_LessThanConst: BIN_CONST(FN(mask)(x < y), FN(mask)(x < y))
#define BIN(EXPR) for (l = 0; l < LB; l++) { REAL x = a[l], y = b[l]; out[l] = EXPR; } break;
    case OP_Add: BIN(x + y)
    case OP_Sub: BIN(x - y)
    case OP_Mul: BIN(x * y)
    case OP_Div: BIN(x / y)
    /* x86 min/max: second operand unless the first compares strictly (Ops.hpp:618-626) */
    case OP_Max: BIN(x > y ? x : y)
    case OP_Min: BIN(x < y ? x : y)
    case OP_Or: BIN(FN(from_bits)(FN(bits)(x) | FN(bits)(y)))
    case OP_And: BIN(FN(from_bits)(FN(bits)(x) & FN(bits)(y)))
    case OP_GreaterThan: BIN(FN(mask)(x > y))
    case OP_GreaterEqual: BIN(FN(mask)(x >= y))
    case OP_LessThan: BIN(FN(mask)(x < y))
    case OP_LessEqual: BIN(FN(mask)(x <= y))
    case OP_Equals: BIN(FN(mask)(x == y))
    case OP_Sqrt: for (l = 0; l < LB; l++) out[l] = SQRT(a[l]); break;
    case OP_Log: for (l = 0; l < LB; l++) out[l] = e->pg->fast_log ? FN(log_fast)(a[l]) : LOGLIB(a[l]); break;
    case OP_Exp: for (l = 0; l < LB; l++) out[l] = FN(exp_fast)(a[l]); break;
    case OP_Abs: for (l = 0; l < LB; l++) out[l] = FN(from_bits)(FN(bits)(a[l]) & ~((UBITS)1 << (sizeof(UBITS) * 8 - 1))); break;
    case OP_Not: for (l = 0; l < LB; l++) out[l] = FN(from_bits)(~FN(bits)(a[l])); break;
    case OP_Sign: /* Ops.hpp:692-701 */
        for (l = 0; l < LB; l++) {
            REAL v = a[l];
            REAL r = (v != v) ? (REAL)NAN : (REAL)1.0;
            if (v < 0) r = -1.0;
            if (v == 0) r = 0.0;
            out[l] = r;
        }
        break;
    case OP_SetInfOrNanToValue: { /* Ops.hpp:715-720 */
        REAL C_ = FN(const_of)(o);
        for (l = 0; l < LB; l++) { REAL d = a[l] - a[l]; out[l] = (d != d) ? C_ : a[l]; }
        break;
    }
    case OP_Select: /* blendv on the sign bit, f32x8.hpp:84-87 */
        for (l = 0; l < LB; l++) out[l] = FN(sign_set)(a[l]) ? b[l] : cc[l];
        break;
    case OP_ForeachBackWindow:
        FN(eval_loop)(c, e, k, row);
        break;
    case OP_IterValue: {
        const REAL *h = FN(ring_at)(c, o->in[1], LB, row, iter);
        memcpy(out, h, sizeof(REAL) * LB);
        break;
    }
    case OP_WindowLoopIndex: { REAL v = (REAL)(o->window - 1 - iter); for (l = 0; l < LB; l++) out[l] = v; break; }
    case OP_BackRef: { /* windowedRef, Ops.hpp:522-535 */
        const REAL *h = FN(ring_at)(c, o->in[0], LB, row, o->window);
        memcpy(out, h, sizeof(REAL) * LB);
        break;
    }
    case OP_FastWindowedSum: { /* Ops.hpp:324-356 */
        REAL *v = c->st[k], *cadd = v + LB, *csub = v + 2 * LB;
        int64_t *nn = c->cnt[k];
        const REAL *old = FN(ring_at)(c, o->in[0], LB, row, o->window);
        for (l = 0; l < LB; l++) {
            REAL ol = old[l], cur = a[l];
            int onan = ol != ol, nnan = cur != cur;
            if (!onan) {
                REAL y = (REAL)0 - ol;
                y = y - csub[l];
                REAL t = v[l] + y;
                csub[l] = (t - v[l]) - y;
                v[l] = t;
            }
            if (!nnan) {
                REAL y = cur - cadd[l];
                REAL t = v[l] + y;
                cadd[l] = (t - v[l]) - y;
                v[l] = t;
            }
            nn[l] = nn[l] - onan + nnan;
            out[l] = nn[l] == 0 ? v[l] : (REAL)NAN;
        }
        break;
    }
    case OP_ExpMovingAvg: { /* Ops.hpp:358-375 */
        REAL *v = c->st[k];
        const REAL w = (REAL)2.0 / (REAL)(o->window + 1);
        const REAL keep = (REAL)1 - w;
        if (o->nin > 1 && row == 0 && !c->rows) for (l = 0; l < LB; l++) v[l] = b[l];
        for (l = 0; l < LB; l++) {
            REAL cur = a[l], old = v[l];
            REAL nv = old * keep;
            REAL add = cur * w;
            nv = nv + add;
            REAL r = (cur != cur) ? old : nv;
            if (old != old) r = cur;
            v[l] = r;
            out[l] = r;
        }
        break;
    }
    case OP_Accumulator:
        memcpy(out, c->st[k], sizeof(REAL) * LB);
        break;
    case OP_SetAccumulator: { /* Ops.hpp:300-322: acc = mask ? v : acc-as-read */
        REAL *acc = c->st[o->in[0]];
        for (l = 0; l < LB; l++) { acc[l] = FN(sign_set)(b[l]) ? cc[l] : a[l]; out[l] = acc[l]; }
        break;
    }
    case OP_WindowedLinearRegression: { /* Ops.hpp:377-423 */
        REAL *is = c->st[k], *xs = is + LB, *x2 = is + 2 * LB, *ys = is + 3 * LB, *y2 = is + 4 * LB, *xy = is + 5 * LB;
        int64_t *nn = c->cnt[k];
        const REAL *old = FN(ring_at)(c, o->in[0], LB, row, o->window);
        const REAL W = (REAL)o->window;
        for (l = 0; l < LB; l++) {
            REAL ol = old[l], cur = a[l];
            int onan = ol != ol, nnan = cur != cur;
            xy[l] = xy[l] - ys[l];
            x2[l] = (x2[l] + is[l]) - (REAL)2.0 * xs[l];
            xs[l] = xs[l] - is[l];
            if (!onan) { is[l] = is[l] - (REAL)1; ys[l] = ys[l] - ol; y2[l] = y2[l] - ol * ol; }
            nn[l] = nn[l] - onan + nnan;
            if (!nnan) {
                is[l] = is[l] + (REAL)1;
                xs[l] = xs[l] + W;
                x2[l] = x2[l] + W * W;
                ys[l] = ys[l] + cur;
                y2[l] = y2[l] + cur * cur;
                xy[l] = xy[l] + W * cur;
            }
            out[l] = 0;
        }
        break;
    }
    case OP_WindowedLinearRegressionRSqaureImpl:
    case OP_WindowedLinearRegressionSlopeImpl:
    case OP_WindowedLinearRegressionResiImpl: { /* Ops.hpp:425-466 */
        int src = o->in[0];
        const REAL *xs = c->st[src] + LB, *x2 = c->st[src] + 2 * LB, *ys = c->st[src] + 3 * LB,
                   *y2 = c->st[src] + 4 * LB, *xy = c->st[src] + 5 * LB;
        const int64_t *nn = c->cnt[src];
        const REAL W = (REAL)sg->ops[src].window;
        for (l = 0; l < LB; l++) {
            REAL N = (REAL)(sg->ops[src].window - nn[l]);
            if (o->kind == OP_WindowedLinearRegressionRSqaureImpl) {
                REAL r1 = N * xy[l] - xs[l] * ys[l];
                r1 = r1 * r1;
                REAL r2 = (N * x2[l] - xs[l] * xs[l]) * (N * y2[l] - ys[l] * ys[l]);
                out[l] = r1 / r2;
            } else {
                REAL slope = (N * xy[l] - xs[l] * ys[l]) / (N * x2[l] - xs[l] * xs[l]);
                if (o->kind == OP_WindowedLinearRegressionSlopeImpl) out[l] = slope;
                else {
                    REAL xm = xs[l] / N, ym = ys[l] / N;
                    REAL interp = ym - slope * xm;
                    out[l] = b[l] - (slope * W + interp);
                }
            }
        }
        break;
    }
    case OP_SkipListState: { /* Ops/Quantile.hpp:15-45 */
        int w = o->window;
        for (l = 0; l < LB; l++) {
            double *vals = c->sk_val[k] + (size_t)l * (w + 1);
            int64_t *idx = c->sk_idx[k] + (size_t)l * (w + 1);
            if (a[l] == a[l]) FN(sk_remove)(vals, idx, &c->sk_n[k][l], (double)a[l]);
            if (b[l] == b[l]) FN(sk_insert)(vals, idx, &c->sk_n[k][l], &c->sk_last[k][l], (double)b[l], (int64_t)i);
        }
        memset(out, 0, sizeof(REAL) * LB);
        break;
    }
    case OP_SkipListQuantile: case OP_SkipListRank: case OP_SkipListMin: case OP_SkipListMax:
    case OP_SkipListArgMin: { /* Ops/Quantile.hpp:84-172 */
        int src = o->in[0];
        int w = sg->ops[src].window;
        REAL q = o->kind == OP_SkipListQuantile ? FN(const_of)(o) : 0;
        for (l = 0; l < LB; l++) {
            double *vals = c->sk_val[src] + (size_t)l * (w + 1);
            int64_t *idx = c->sk_idx[src] + (size_t)l * (w + 1);
            int n = c->sk_n[src][l];
            if (n != w) { out[l] = NAN; continue; }
            switch (o->kind) {
            case OP_SkipListQuantile: {
                REAL f = q * (REAL)(n - 1);
                int ii = (int)f;
                if ((REAL)ii == f) out[l] = (REAL)vals[ii];
                else {
                    double lo = vals[ii], hi = vals[ii + 1];
                    out[l] = (REAL)(lo + (hi - lo) * (double)(f - (REAL)ii));
                }
                break;
            }
            case OP_SkipListRank: {
                double cur = (double)b[l];
                int lt = 0;
                while (lt < n && vals[lt] < cur) lt++;
                double rank = c->sk_last[src][l] + 1;
                double rank_min = lt + 1;
                rank = (((rank * (rank + 1) / 2) - ((rank_min - 1) * rank_min / 2)) / (rank - rank_min + 1));
                out[l] = (REAL)rank;
                break;
            }
            case OP_SkipListMin: out[l] = (REAL)vals[0]; break;
            case OP_SkipListMax: out[l] = (REAL)vals[w - 1]; break;
            default: out[l] = (REAL)(int64_t)(idx[0] + n - i); break;
            }
        }
        break;
    }
    case OP_ReduceAdd: case OP_ReduceMul: case OP_ReduceMin: case OP_ReduceMax:
    case OP_ReduceArgMax: case OP_ReduceArgMin: case OP_ReduceRank: case OP_ReduceDecayLinear:
        /* computed when the loop ran */
        break;
    default:
        fprintf(stderr, "kun_oracle: unsupported op %d\n", o->kind);
        abort();
    }
    if (c->ring[k]) FN(ring_push)(c, k, LB, out, row);
}

/* Runs window loop `L` for the current row: body ops for every offset from
 * the oldest (window-1) to the newest (0), then the attached reductions
 * (Ops.hpp:468-616; loop order CodegenCpp.py:246). Loops merged by
 * merge_loops are evaluated as independent full loops, which gives every
 * reduction the same oldest-to-newest sequence. */
static void FN(eval_loop)(FN(Ctx) *c, const FN(StepEnv) *e, int L, long row) {
    const Stage *sg = e->sg;
    int LB = e->LB, l;
    int w = sg->ops[L].window;
    int nred = 0, reds[512];
    for (int k = 0; k < sg->nops && nred < 512; k++)
        if (sg->ops[k].loop == L) reds[nred++] = k;
    /* reduction state: s0..s3 per lane, in the reduction's st slots */
    for (int r = 0; r < nred; r++) {
        int k = reds[r];
        const Op *o = &sg->ops[k];
        if (!c->st[k]) c->st[k] = (REAL *)calloc((size_t)8 * LB, sizeof(REAL));
        REAL *s0 = c->st[k], *s1 = s0 + LB, *s2 = s0 + 2 * LB;
        for (l = 0; l < LB; l++) {
            switch (o->kind) {
            case OP_ReduceAdd: s0[l] = 0; s1[l] = 0; break;
            case OP_ReduceMul: s0[l] = 1; break;
            case OP_ReduceMin: s0[l] = INFINITY; break;
            case OP_ReduceMax: s0[l] = -INFINITY; break;
            case OP_ReduceArgMax: s0[l] = -INFINITY; s1[l] = 0; break;
            case OP_ReduceArgMin: s0[l] = INFINITY; s1[l] = 0; break;
            case OP_ReduceRank: {
                /* The reference constructs the rank state from its current
                 * value (`ReduceRank<T, N> v{cur}`, CodegenCpp.py:246-270);
                 * when that value is itself a ReduceRank (directly or
                 * through Output / WindowedTempOutput / ReturnFirstValue
                 * copies), C++ picks the copy constructor: the new rank
                 * starts from the other's comparison value and counts
                 * (Ops.hpp:598-616). */
                int src = o->in[1];
                while (sg->ops[src].kind == OP_Output || sg->ops[src].kind == OP_WindowedTempOutput ||
                       sg->ops[src].kind == OP_ReturnFirstValue)
                    src = sg->ops[src].in[0];
                if (sg->ops[src].kind == OP_ReduceRank && c->st[src]) {
                    const REAL *p0 = c->st[src];
                    s0[l] = p0[l]; s1[l] = p0[LB + l]; s2[l] = p0[2 * LB + l];
                } else {
                    s0[l] = V(o->in[1])[l]; s1[l] = 0; s2[l] = 0;
                }
                break;
            }
            case OP_ReduceDecayLinear:
                s0[l] = (REAL)(1.0 / ((1.0 + o->window) * o->window / 2)); s1[l] = 0; break;
            default: break;
            }
        }
    }
    for (int iter = w - 1; iter >= 0; iter--) {
        for (int k = L + 1; k < sg->nops; k++)
            if (sg->ops[k].parent == L) FN(eval_op)(c, e, k, row, row, iter);
        for (int r = 0; r < nred; r++) {
            int k = reds[r];
            const Op *o = &sg->ops[k];
            const REAL *in = V(o->in[0]);
            REAL *s0 = c->st[k], *s1 = s0 + LB, *s2 = s0 + 2 * LB;
            for (l = 0; l < LB; l++) {
                REAL x = in[l];
                switch (o->kind) {
                case OP_ReduceAdd: { /* kahanAdd, Ops.hpp:289-298 */
                    REAL y = x - s1[l];
                    REAL t = s0[l] + y;
                    s1[l] = (t - s0[l]) - y;
                    s0[l] = t;
                    break;
                }
                case OP_ReduceMul: s0[l] = s0[l] * x; break;
                case OP_ReduceMin: { REAL v = s0[l]; s0[l] = (v != v || x != x) ? (REAL)NAN : (v < x ? v : x); break; }
                case OP_ReduceMax: { REAL v = s0[l]; s0[l] = (v != v || x != x) ? (REAL)NAN : (v > x ? v : x); break; }
                case OP_ReduceArgMax: case OP_ReduceArgMin: {
                    REAL v = s0[l];
                    int isn = (v != v) || (x != x);
                    int cmp = o->kind == OP_ReduceArgMax ? (v < x) : (v > x);
                    v = cmp ? x : v;
                    s0[l] = isn ? (REAL)NAN : v;
                    REAL id = cmp ? (REAL)iter : s1[l];
                    s1[l] = isn ? (REAL)NAN : id;
                    break;
                }
                case OP_ReduceRank: {
                    REAL v = s0[l];
                    int isn = (v != v) || (x != x);
                    if (isn) s1[l] = NAN;
                    if (x < v) s1[l] = s1[l] + (REAL)1.0;
                    if (x == v) s2[l] = s2[l] + (REAL)1.0;
                    break;
                }
                case OP_ReduceDecayLinear: {
                    const REAL step = (REAL)(1.0 / ((1.0 + o->window) * o->window / 2));
                    s1[l] = FMA(x, s0[l], s1[l]);
                    s0[l] = s0[l] + step;
                    break;
                }
                default: break;
                }
            }
        }
    }
    for (int r = 0; r < nred; r++) {
        int k = reds[r];
        const Op *o = &sg->ops[k];
        REAL *s0 = c->st[k], *s1 = s0 + LB, *s2 = s0 + 2 * LB;
        REAL *out = V(k);
        for (l = 0; l < LB; l++) {
            switch (o->kind) {
            case OP_ReduceAdd: case OP_ReduceMul: case OP_ReduceMin: case OP_ReduceMax: out[l] = s0[l]; break;
            case OP_ReduceArgMax: case OP_ReduceArgMin: out[l] = s1[l]; break;
            case OP_ReduceRank: out[l] = s1[l] + (s2[l] + (REAL)1.0) / (REAL)2.0; break;
            case OP_ReduceDecayLinear: out[l] = s1[l]; break;
            default: break;
            }
        }
        if (c->ring[k]) FN(ring_push)(c, k, LB, out, row);
    }
}

static void FN(step)(FN(Ctx) *c, const FN(StepEnv) *e, long row, long i) {
    const Stage *sg = e->sg;
    for (int k = 0; k < sg->nops; k++) {
        const Op *o = &sg->ops[k];
        if (o->parent >= 0 || o->loop >= 0) continue;
        FN(eval_op)(c, e, k, row, i, 0);
    }
    c->rows++;
}

#undef V

/* ---- cross-sectional rows (Rank.hpp:11-57, Scale.hpp:13-46) ---------- */
static int FN(cmp_real)(const void *x, const void *y) {
    REAL a = *(const REAL *)x, b = *(const REAL *)y;
    return a < b ? -1 : (b < a ? 1 : 0);
}

static void FN(rank_row)(const REAL *in, REAL *out, long S, REAL *scratch) {
    long n = 0;
    for (long s = 0; s < S; s++) if (in[s] == in[s]) scratch[n++] = in[s];
    qsort(scratch, n, sizeof(REAL), FN(cmp_real));
    for (long s = 0; s < S; s++) {
        REAL v = in[s];
        if (v != v) { out[s] = NAN; continue; }
        long lo = 0, hi = n;
        while (lo < hi) { long m = (lo + hi) / 2; if (scratch[m] < v) lo = m + 1; else hi = m; }
        long st = lo;
        hi = n;
        while (lo < hi) { long m = (lo + hi) / 2; if (v < scratch[m]) hi = m; else lo = m + 1; }
        long en = lo;
        long sum = (st + en + 1) * (en - st) / 2;
        out[s] = (REAL)sum / (REAL)(en - st) / (REAL)n;
    }
}

static void FN(scale_row)(const REAL *in, REAL *out, long S) {
    REAL sum = 0;
    for (long s = 0; s < S; s++) if (in[s] == in[s]) sum += FN(from_bits)(FN(bits)(in[s]) & ~((UBITS)1 << (sizeof(UBITS) * 8 - 1)));
    for (long s = 0; s < S; s++) out[s] = (in[s] == 0 && sum == 0) ? (REAL)NAN : in[s] / sum;
}

/* ---- aggregrate: per stock, rows split into groups where the label
 * changes; per group sum (kahanAdd), min, max (NaN if any NaN), first,
 * last, count (every row), mean = sum / count -- Aggregration.cpp:11-105.
 * outs[k] for k = sum, min, max, first, last, count, mean (NULL: not
 * wanted) are [groups][S]; rows start .. start + L - 1 of in ([T][S]) and
 * labels ([T]) are read. */
void FN(kun_oracle_aggregate)(const REAL *in, const REAL *labels, REAL *const *outs, long S, long T, long start,
                              long L) {
    (void)T;
    if (L <= 0) return;
    for (long s = 0; s < S; s++) {
        REAL sum = 0, comp = 0, mn = INFINITY, mx = -INFINITY, first = in[start * S + s], last = 0, count = 0;
        REAL last_label = labels[start];
        long g = 0;
        for (long i = 0; i < L; i++) {
            REAL label = labels[start + i];
            REAL cur = in[(start + i) * S + s];
            if (label != last_label) {
                REAL vals[7] = {sum, mn, mx, first, last, count, sum / count};
                for (int k = 0; k < 7; k++) if (outs[k]) outs[k][g * S + s] = vals[k];
                g++;
                first = cur;
                sum = 0; comp = 0; mn = INFINITY; mx = -INFINITY; count = 0;
            }
            count = count + (REAL)1;
            last = cur;
            last_label = label;
            mx = (mx != mx || cur != cur) ? (REAL)NAN : (mx > cur ? mx : cur);
            mn = (mn != mn || cur != cur) ? (REAL)NAN : (mn < cur ? mn : cur);
            REAL y = cur - comp;
            REAL t = sum + y;
            comp = (t - sum) - y;
            sum = t;
        }
        REAL vals[7] = {sum, mn, mx, first, last, count, sum / count};
        for (int k = 0; k < 7; k++) if (outs[k]) outs[k][g * S + s] = vals[k];
    }
}
