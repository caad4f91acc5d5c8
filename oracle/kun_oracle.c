/* kun_oracle: CPU restatement of the KunQuant factor-pipeline semantics.
 *
 * TEST INFRASTRUCTURE ONLY. Used by tests/, by __graft_entry__.smoke() and
 * by bench.py's cpu_baseline leg as the checker and the timed CPU baseline;
 * the product (kunquant_amd, libkunhip.so, generated device modules) never
 * links, loads or calls it.
 *
 * It interprets the KUNPROG text of a compiled module (buffer table, stage
 * DAG and per-stage op lists produced by kunquant_amd.codegen.program) with
 * the arithmetic of the reference runtime:
 *   - stage execution, buffers and the no-history-before-start rule:
 *       cpp/Kun/Runtime.cpp:311-352, Ops.hpp:39-210, CodegenCpp.py:141-307
 *   - primitives: cpp/Kun/Ops.hpp, Ops/Quantile.hpp, Rank.hpp, Scale.hpp
 *   - lane semantics: cpp/KunSIMD/cpu/x86/f32x8.hpp, f64x4.hpp
 *   - LogFast / Exp: cpp/KunSIMD/cpu/Math.hpp
 * Build: -O2 -ffp-contract=off (no FMA except the explicit fma() calls that
 * mirror the reference's sc_fmadd).
 */
#include <math.h>
#include <pthread.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "../kunquant_amd/csrc/kun_logtables.h"

enum {
    OP_Input, OP_Output, OP_WindowedTempOutput, OP_ConstantOp, OP_AddConst, OP_SubConst,
    OP_MulConst, OP_DivConst, OP_GreaterThanConst, OP_LessThanConst, OP_Add, OP_Sub, OP_Mul,
    OP_Div, OP_Max, OP_Min, OP_Or, OP_And, OP_GreaterThan, OP_GreaterEqual, OP_LessThan,
    OP_LessEqual, OP_Equals, OP_Sqrt, OP_Log, OP_Abs, OP_Sign, OP_Not, OP_Exp,
    OP_SetInfOrNanToValue, OP_Select, OP_ForeachBackWindow, OP_IterValue, OP_WindowLoopIndex,
    OP_ReduceAdd, OP_ReduceMul, OP_ReduceMin, OP_ReduceMax, OP_ReduceArgMax, OP_ReduceArgMin,
    OP_ReduceRank, OP_ReduceDecayLinear, OP_BackRef, OP_FastWindowedSum, OP_ExpMovingAvg,
    OP_Accumulator, OP_SetAccumulator, OP_ReturnFirstValue, OP_WindowedLinearRegression,
    OP_WindowedLinearRegressionRSqaureImpl, OP_WindowedLinearRegressionSlopeImpl,
    OP_WindowedLinearRegressionResiImpl, OP_SkipListState, OP_SkipListQuantile,
    OP_SkipListRank, OP_SkipListMin, OP_SkipListMax, OP_SkipListArgMin, OP_COUNT
};

static const char *OP_NAMES[OP_COUNT] = {
    "Input", "Output", "WindowedTempOutput", "ConstantOp", "AddConst", "SubConst",
    "MulConst", "DivConst", "GreaterThanConst", "LessThanConst", "Add", "Sub", "Mul",
    "Div", "Max", "Min", "Or", "And", "GreaterThan", "GreaterEqual", "LessThan",
    "LessEqual", "Equals", "Sqrt", "Log", "Abs", "Sign", "Not", "Exp",
    "SetInfOrNanToValue", "Select", "ForeachBackWindow", "IterValue", "WindowLoopIndex",
    "ReduceAdd", "ReduceMul", "ReduceMin", "ReduceMax", "ReduceArgMax", "ReduceArgMin",
    "ReduceRank", "ReduceDecayLinear", "BackRef", "FastWindowedSum", "ExpMovingAvg",
    "Accumulator", "SetAccumulator", "ReturnFirstValue", "WindowedLinearRegression",
    "WindowedLinearRegressionRSqaureImpl", "WindowedLinearRegressionSlopeImpl",
    "WindowedLinearRegressionResiImpl", "SkipListState", "SkipListQuantile",
    "SkipListRank", "SkipListMin", "SkipListMax", "SkipListArgMin"};

typedef struct {
    int kind, parent, loop, nin, in[4], buf, window, swap, single;
    char value[64];
} Op;

typedef struct {
    int id, kind, nops, nin, nout, ndep;
    int in[256], out[256], dep[256];
    Op *ops;
} Stage;

typedef struct {
    int id, kind, window, unreliable;
    char name[256];
} Buf;

typedef struct {
    char name[256];
    int dtype; /* 0 float, 1 double */
    int in_layout, out_layout, blocking, aligned, fast_log;
    int nbuf, nstage;
    Buf *bufs;
    Stage *stages;
    int *order; /* topological stage order */
} Prog;

static void prog_free(Prog *p) {
    if (!p) return;
    for (int i = 0; i < p->nstage; i++) free(p->stages[i].ops);
    free(p->stages); free(p->bufs); free(p->order); free(p);
}

static int op_code(const char *name) {
    for (int i = 0; i < OP_COUNT; i++) if (!strcmp(OP_NAMES[i], name)) return i;
    return -1;
}

static const char *next_line(const char *p, char *line, size_t cap) {
    if (!*p) return NULL;
    size_t n = 0;
    while (*p && *p != '\n') { if (n + 1 < cap) line[n++] = *p; p++; }
    line[n] = 0;
    if (*p == '\n') p++;
    return p;
}

void *kun_oracle_load(const char *text, char *err, int errlen) {
    Prog *pg = (Prog *)calloc(1, sizeof(Prog));
    pg->fast_log = 1;
    char line[8192];
    const char *p = text;
    int cur_stage = -1, cur_op = 0;
#define FAIL(msg) do { snprintf(err, errlen, "%s: %s", msg, line); prog_free(pg); return NULL; } while (0)
    while ((p = next_line(p, line, sizeof line))) {
        char tag[64];
        if (sscanf(line, "%63s", tag) != 1) continue;
        if (!strcmp(tag, "KUNPROG") || !strcmp(tag, "end")) continue;
        if (!strcmp(tag, "module")) { sscanf(line, "module %255s", pg->name); }
        else if (!strcmp(tag, "dtype")) { char d[32]; sscanf(line, "dtype %31s", d); pg->dtype = !strcmp(d, "double"); }
        else if (!strcmp(tag, "layout")) sscanf(line, "layout %d %d", &pg->in_layout, &pg->out_layout);
        else if (!strcmp(tag, "blocking")) sscanf(line, "blocking %d", &pg->blocking);
        else if (!strcmp(tag, "aligned")) sscanf(line, "aligned %d", &pg->aligned);
        else if (!strcmp(tag, "fastlog")) sscanf(line, "fastlog %d", &pg->fast_log);
        else if (!strcmp(tag, "buffers")) {
            sscanf(line, "buffers %d", &pg->nbuf);
            pg->bufs = (Buf *)calloc(pg->nbuf, sizeof(Buf));
        } else if (!strcmp(tag, "B")) {
            Buf b;
            memset(&b, 0, sizeof b);
            if (sscanf(line, "B %d %d %d %d %255s", &b.id, &b.kind, &b.window, &b.unreliable, b.name) != 5) FAIL("bad buffer line");
            if (b.id < 0 || b.id >= pg->nbuf) FAIL("bad buffer id");
            pg->bufs[b.id] = b;
        } else if (!strcmp(tag, "stages")) {
            sscanf(line, "stages %d", &pg->nstage);
            pg->stages = (Stage *)calloc(pg->nstage, sizeof(Stage));
        } else if (!strcmp(tag, "S")) {
            Stage *s;
            char *q = line + 1;
            int id = (int)strtol(q, &q, 10);
            if (id < 0 || id >= pg->nstage) FAIL("bad stage id");
            s = &pg->stages[id];
            s->id = id;
            s->kind = (int)strtol(q, &q, 10);
            if (s->kind < 0 || s->kind > 3) FAIL("stage kind the oracle cannot run (user cross-sectional body)");
            s->nops = (int)strtol(q, &q, 10);
            s->nin = (int)strtol(q, &q, 10);
            if (s->nin > 256) FAIL("too many inputs");
            for (int i = 0; i < s->nin; i++) s->in[i] = (int)strtol(q, &q, 10);
            s->nout = (int)strtol(q, &q, 10);
            if (s->nout > 256) FAIL("too many outputs");
            for (int i = 0; i < s->nout; i++) s->out[i] = (int)strtol(q, &q, 10);
            s->ndep = (int)strtol(q, &q, 10);
            if (s->ndep > 256) FAIL("too many deps");
            for (int i = 0; i < s->ndep; i++) s->dep[i] = (int)strtol(q, &q, 10);
            s->ops = (Op *)calloc(s->nops > 0 ? s->nops : 1, sizeof(Op));
            cur_stage = id;
            cur_op = 0;
        } else if (!strcmp(tag, "O")) {
            if (cur_stage < 0 || cur_op >= pg->stages[cur_stage].nops) FAIL("op outside stage");
            Op *o = &pg->stages[cur_stage].ops[cur_op++];
            char name[128];
            char *q = line + 1;
            while (*q == ' ') q++;
            int n = 0;
            sscanf(q, "%127s%n", name, &n);
            q += n;
            o->kind = op_code(name);
            if (o->kind < 0) FAIL("unknown op");
            o->parent = (int)strtol(q, &q, 10);
            o->loop = (int)strtol(q, &q, 10);
            o->nin = (int)strtol(q, &q, 10);
            if (o->nin > 4) FAIL("too many op inputs");
            for (int i = 0; i < o->nin; i++) o->in[i] = (int)strtol(q, &q, 10);
            o->buf = (int)strtol(q, &q, 10);
            o->window = (int)strtol(q, &q, 10);
            o->swap = (int)strtol(q, &q, 10);
            o->single = (int)strtol(q, &q, 10);
            sscanf(q, "%63s", o->value);
        }
    }
#undef FAIL
    /* topological order of stages (Kahn) */
    pg->order = (int *)malloc(sizeof(int) * (pg->nstage + 1));
    int *indeg = (int *)calloc(pg->nstage + 1, sizeof(int));
    int head = 0, tail = 0;
    for (int i = 0; i < pg->nstage; i++) indeg[i] = pg->stages[i].ndep;
    for (int i = 0; i < pg->nstage; i++) if (!indeg[i]) pg->order[tail++] = i;
    while (head < tail) {
        int s = pg->order[head++];
        for (int j = 0; j < pg->nstage; j++)
            for (int d = 0; d < pg->stages[j].ndep; d++)
                if (pg->stages[j].dep[d] == s && --indeg[j] == 0) pg->order[tail++] = j;
    }
    free(indeg);
    if (tail != pg->nstage) { snprintf(err, errlen, "stage graph has a cycle"); prog_free(pg); return NULL; }
    return pg;
}

void kun_oracle_free(void *p) { prog_free((Prog *)p); }
int kun_oracle_num_buffers(void *p) { return ((Prog *)p)->nbuf; }

/* ---- element-type instantiations ---------------------------------------- */
#define REAL float
#define UBITS uint32_t
#define SBITS int32_t
#define FN(x) x##_f32
#define MANT 23
#define EXPB 8
#define LOG_R KUN_LOG_R_F32
#define LOG_LR KUN_LOG_LR_F32
#define FMA fmaf
#define FLOOR floorf
#define SQRT sqrtf
#define LOGLIB logf
#define PARSE strtof_
static float strtof_(const char *s) { return strtof(s, NULL); }
#include "kun_oracle_impl.h"
#undef REAL
#undef UBITS
#undef SBITS
#undef FN
#undef MANT
#undef EXPB
#undef LOG_R
#undef LOG_LR
#undef FMA
#undef FLOOR
#undef SQRT
#undef LOGLIB
#undef PARSE

#define REAL double
#define UBITS uint64_t
#define SBITS int64_t
#define FN(x) x##_f64
#define MANT 52
#define EXPB 11
#define LOG_R KUN_LOG_R_F64
#define LOG_LR KUN_LOG_LR_F64
#define FMA fma
#define FLOOR floor
#define SQRT sqrt
#define LOGLIB log
#define PARSE strtod_
static double strtod_(const char *s) { return strtod(s, NULL); }
#include "kun_oracle_impl.h"

/* ---- batch driver -------------------------------------------------------- */
#define LANE_BLOCK 64

typedef struct {
    Prog *pg;
    const Stage *sg;
    void *views; /* BufView_f32/f64 array */
    long S, start, L;
    long item0, item1; /* lane blocks or rows */
} Job;

static void *ts_worker_f32(void *arg) {
    Job *j = (Job *)arg;
    for (long blk = j->item0; blk < j->item1; blk++) {
        Ctx_f32 *c = ctx_new_f32(j->sg, LANE_BLOCK, NULL);
        StepEnv_f32 e = {j->pg, j->sg, (BufView_f32 *)j->views, j->S, j->start, blk * LANE_BLOCK, LANE_BLOCK, 0};
        e.nl = (int)((j->S - e.lane0) < LANE_BLOCK ? (j->S - e.lane0) : LANE_BLOCK);
        for (long i = 0; i < j->L; i++) step_f32(c, &e, i, i);
        ctx_free_f32(c, j->sg);
    }
    return NULL;
}

static void *ts_worker_f64(void *arg) {
    Job *j = (Job *)arg;
    for (long blk = j->item0; blk < j->item1; blk++) {
        Ctx_f64 *c = ctx_new_f64(j->sg, LANE_BLOCK, NULL);
        StepEnv_f64 e = {j->pg, j->sg, (BufView_f64 *)j->views, j->S, j->start, blk * LANE_BLOCK, LANE_BLOCK, 0};
        e.nl = (int)((j->S - e.lane0) < LANE_BLOCK ? (j->S - e.lane0) : LANE_BLOCK);
        for (long i = 0; i < j->L; i++) step_f64(c, &e, i, i);
        ctx_free_f64(c, j->sg);
    }
    return NULL;
}

/* cross-sectional rows: gather row(s), compute, scatter. Stage kinds:
 * 1 RankStocks, 2 ScaleStocks, 3 DiffWithWeightedSum (MiscOp.py:227-249:
 * acc = sum over stocks in order of in0*in1, each product and addition
 * rounded in T; out = in0 - acc). */
#define CS_WORKER(SFX, REALT)                                                                  \
    static void gather_##SFX(Job *j, int b, long t, REALT *row) {                              \
        BufView_##SFX *v = (BufView_##SFX *)j->views;                                          \
        const Buf *bi = &j->pg->bufs[b];                                                       \
        long r = (bi->kind == 0 ? j->start : 0) + t - v[b].base;                               \
        for (long s = 0; s < j->S; s++)                                                        \
            row[s] = ((REALT *)v[b].ptr)[buf_index_##SFX(&v[b], r, s, j->S, j->pg->blocking)];   \
    }                                                                                          \
    static void *cs_worker_##SFX(void *arg) {                                                  \
        Job *j = (Job *)arg;                                                                   \
        BufView_##SFX *v = (BufView_##SFX *)j->views;                                          \
        int ob = j->sg->out[0];                                                                \
        REALT *row = (REALT *)malloc(sizeof(REALT) * j->S * 4);                                \
        REALT *res = row + j->S, *tmp = row + 2 * j->S, *row2 = row + 3 * j->S;                \
        for (long t = j->item0; t < j->item1; t++) {                                           \
            gather_##SFX(j, j->sg->in[0], t, row);                                             \
            if (j->sg->kind == 1) rank_row_##SFX(row, res, j->S, tmp);                         \
            else if (j->sg->kind == 2) scale_row_##SFX(row, res, j->S);                        \
            else {                                                                             \
                gather_##SFX(j, j->sg->in[1], t, row2);                                        \
                REALT acc = 0;                                                                 \
                for (long s = 0; s < j->S; s++) {                                              \
                    REALT prod = row[s] * row2[s];                                             \
                    acc = acc + prod;                                                          \
                }                                                                              \
                for (long s = 0; s < j->S; s++) res[s] = row[s] - acc;                         \
            }                                                                                  \
            for (int o = 0; o < j->sg->nout; o++) {                                            \
                ob = j->sg->out[o];                                                            \
                for (long s = 0; s < j->S; s++)                                                \
                    ((REALT *)v[ob].ptr)[buf_index_##SFX(&v[ob], t, s, j->S, j->pg->blocking)] = res[s]; \
            }                                                                                  \
        }                                                                                      \
        free(row);                                                                             \
        return NULL;                                                                           \
    }
CS_WORKER(f32, float)
CS_WORKER(f64, double)

static void run_parallel(void *(*fn)(void *), Job proto, long nitems, int nthreads) {
    if (nthreads < 1) nthreads = 1;
    if (nthreads > nitems) nthreads = (int)(nitems > 0 ? nitems : 1);
    pthread_t th[256];
    Job jobs[256];
    if (nthreads > 256) nthreads = 256;
    long per = (nitems + nthreads - 1) / nthreads;
    for (int t = 0; t < nthreads; t++) {
        jobs[t] = proto;
        jobs[t].item0 = t * per;
        jobs[t].item1 = (t + 1) * per < nitems ? (t + 1) * per : nitems;
        if (t == 0) continue;
        pthread_create(&th[t], NULL, fn, &jobs[t]);
    }
    fn(&jobs[0]);
    for (int t = 1; t < nthreads; t++) pthread_join(th[t], NULL);
}

/* Runs the whole module in batch mode. bufs[id] must point at every user
 * INPUT/OUTPUT buffer (layouts as declared by the module); temporaries are
 * allocated here. Returns 0 on success. */
int kun_oracle_run(void *prog, void **bufs, long S, long T, long start, long L, int nthreads,
                   char *err, int errlen) {
    Prog *pg = (Prog *)prog;
    size_t esz = pg->dtype ? sizeof(double) : sizeof(float);
    /* one view per buffer; the two instantiations share the layout */
    BufView_f64 *views = (BufView_f64 *)calloc(pg->nbuf, sizeof(BufView_f64));
    void **owned = (void **)calloc(pg->nbuf, sizeof(void *));
    for (int b = 0; b < pg->nbuf; b++) {
        const Buf *bi = &pg->bufs[b];
        if (bi->kind == 2) {
            /* temporaries: [L][S], caller-provided when the caller wants them */
            if (!bufs[b]) owned[b] = calloc((size_t)S * L, esz);
            views[b].ptr = bufs[b] ? bufs[b] : owned[b];
            views[b].layout = 1;
            views[b].num_time = L;
        } else {
            if (!bufs[b]) { snprintf(err, errlen, "missing buffer %s", bi->name); free(views); free(owned); return 1; }
            views[b].ptr = bufs[b];
            views[b].layout = bi->kind == 0 ? pg->in_layout : pg->out_layout;
            views[b].num_time = bi->kind == 0 ? T : L;
        }
        views[b].base = 0;
    }
    for (int oi = 0; oi < pg->nstage; oi++) {
        const Stage *sg = &pg->stages[pg->order[oi]];
        Job proto = {pg, sg, views, S, start, L, 0, 0};
        if (sg->kind == 0) {
            long nblk = (S + LANE_BLOCK - 1) / LANE_BLOCK;
            run_parallel(pg->dtype ? ts_worker_f64 : ts_worker_f32, proto, nblk, nthreads);
        } else {
            run_parallel(pg->dtype ? cs_worker_f64 : cs_worker_f32, proto, L, nthreads);
        }
    }
    for (int b = 0; b < pg->nbuf; b++) free(owned[b]);
    free(owned);
    free(views);
    return 0;
}

/* ---- stream mode ----------------------------------------------------------
 * One persistent context per time-series stage holding the history rings and
 * the carried op state (StreamContext, Runtime.cpp:424-546); every buffer is a
 * single current row. run() advances every stage by one row. */
typedef struct {
    Prog *pg;
    long S;
    void **rows;     /* per buffer: current row [S] */
    void **ctx;      /* per stage: Ctx_f32/Ctx_f64 over all lanes */
    BufView_f64 *views;
    long tick;
} Stream;

void *kun_oracle_stream_create(void *prog, long S) {
    Prog *pg = (Prog *)prog;
    Stream *st = (Stream *)calloc(1, sizeof(Stream));
    size_t esz = pg->dtype ? sizeof(double) : sizeof(float);
    st->pg = pg;
    st->S = S;
    st->rows = (void **)calloc(pg->nbuf, sizeof(void *));
    st->views = (BufView_f64 *)calloc(pg->nbuf, sizeof(BufView_f64));
    for (int b = 0; b < pg->nbuf; b++) {
        st->rows[b] = malloc(esz * S);
        if (pg->dtype) for (long s = 0; s < S; s++) ((double *)st->rows[b])[s] = NAN;
        else for (long s = 0; s < S; s++) ((float *)st->rows[b])[s] = NAN;
        st->views[b].ptr = st->rows[b];
        st->views[b].layout = 2;
        st->views[b].num_time = 1;
    }
    st->ctx = (void **)calloc(pg->nstage, sizeof(void *));
    for (int i = 0; i < pg->nstage; i++) {
        if (pg->stages[i].kind != 0) continue;
        st->ctx[i] = pg->dtype ? (void *)ctx_new_f64(&pg->stages[i], (int)S, NULL)
                               : (void *)ctx_new_f32(&pg->stages[i], (int)S, NULL);
    }
    return st;
}

void *kun_oracle_stream_row(void *sp, int buf) { return ((Stream *)sp)->rows[buf]; }

void kun_oracle_stream_run(void *sp) {
    Stream *st = (Stream *)sp;
    Prog *pg = st->pg;
    for (int oi = 0; oi < pg->nstage; oi++) {
        int si = pg->order[oi];
        const Stage *sg = &pg->stages[si];
        if (sg->kind == 0) {
            if (pg->dtype) {
                StepEnv_f64 e = {pg, sg, st->views, st->S, 0, 0, (int)st->S, (int)st->S};
                step_f64((Ctx_f64 *)st->ctx[si], &e, st->tick, 0);
            } else {
                StepEnv_f32 e = {pg, sg, (BufView_f32 *)st->views, st->S, 0, 0, (int)st->S, (int)st->S};
                step_f32((Ctx_f32 *)st->ctx[si], &e, st->tick, 0);
            }
        } else {
            Job j = {pg, sg, st->views, st->S, 0, 1, 0, 1};
            if (pg->dtype) cs_worker_f64(&j); else cs_worker_f32(&j);
        }
    }
    st->tick++;
}

void kun_oracle_stream_free(void *sp) {
    Stream *st = (Stream *)sp;
    Prog *pg = st->pg;
    for (int i = 0; i < pg->nstage; i++) {
        if (!st->ctx[i]) continue;
        if (pg->dtype) ctx_free_f64((Ctx_f64 *)st->ctx[i], &pg->stages[i]);
        else ctx_free_f32((Ctx_f32 *)st->ctx[i], &pg->stages[i]);
    }
    for (int b = 0; b < pg->nbuf; b++) free(st->rows[b]);
    free(st->rows); free(st->ctx); free(st->views); free(st);
}

/* ---- corrWith / rank corrWith (float), CorrWith.hpp:11-121 ------------
 * Per row t in [start, start + L): sums over the stocks in stock order of
 * x, y, x*x, y*y, x*y (no NaN skipping), each divided by S, then
 * (XY - X*Y) / (sqrt(X2 - X*X) * sqrt(Y2 - Y*Y)) into out[t - start]. With
 * rank_inputs, x is first the row's rank (RankStocks' formula, NaN stays
 * NaN). Layout of x and y: TS [T][S], or STs [S/blk][T][blk]. */
void kun_oracle_corrwith(const float *x, const float *y, float *out, long S, long T, long start, long L, int sts,
                         int blk, int rank_inputs) {
    float *row = (float *)malloc(sizeof(float) * S * 3);
    float *rk = row + S, *tmp = row + 2 * S;
    for (long t = start; t < start + L; t++) {
        for (long s = 0; s < S; s++) row[s] = x[sts ? ((s / blk) * T + t) * blk + s % blk : t * S + s];
        const float *xv = row;
        if (rank_inputs) {
            rank_row_f32(row, rk, S, tmp);
            xv = rk;
        }
        float XY = 0, X = 0, Y = 0, X2 = 0, Y2 = 0;
        for (long s = 0; s < S; s++) {
            float a = xv[s];
            float b = y[sts ? ((s / blk) * T + t) * blk + s % blk : t * S + s];
            X += a;
            Y += b;
            X2 += a * a;
            Y2 += b * b;
            XY += a * b;
        }
        float n = (float)S;
        XY /= n;
        X /= n;
        Y /= n;
        X2 /= n;
        Y2 /= n;
        out[t - start] = (XY - X * Y) / (sqrtf(X2 - X * X) * sqrtf(Y2 - Y * Y));
    }
    free(row);
}
