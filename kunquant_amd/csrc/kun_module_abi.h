/* Descriptor ABI between generated device modules and the device runtime.
 *
 * A compiled module library exports one `const KunModuleDesc kun_module_<name>`
 * per module; the runtime finds it with dlsym (the MI355X counterpart of the
 * reference's `KUN_EXPORT Module <name>` symbol, Driver.py:324-337,
 * Module.cpp:36-38). Stage kernels are referenced by their host stubs and
 * launched through hipLaunchKernel. */
#ifndef KUN_MODULE_ABI_H
#define KUN_MODULE_ABI_H

#define KUN_MODULE_ABI_VERSION 0x4B554E09u

enum { KUN_BUF_INPUT = 0, KUN_BUF_OUTPUT = 1, KUN_BUF_TEMP = 2 };
enum { KUN_STAGE_TS = 0, KUN_STAGE_RANK = 1, KUN_STAGE_SCALE = 2, KUN_STAGE_GENERIC = 3 };
enum { KUN_LAYOUT_STS = 0, KUN_LAYOUT_TS = 1, KUN_LAYOUT_STREAM = 2 };
enum { KUN_DTYPE_FLOAT = 0, KUN_DTYPE_DOUBLE = 1 };
enum { KUN_KERNEL_PAR = 0, KUN_KERNEL_SEQ = 1 };

enum { KUN_PHASE_TS = 0, KUN_PHASE_CS = 1 };

/* One launch of a session pool's tick (STREAM modules): KUN_PHASE_TS, a
 * kernel taking kun::PoolArgs in which every stock's lane runs the tick
 * bodies of `stages` in order; KUN_PHASE_CS, the cross-sectional `stages`
 * over every session's stock range (kernel is null). */
typedef struct {
    int kind;
    const void *kernel;
    const int *stages;
    int nstages;
} KunPhaseDesc;

typedef struct {
    const char *name;
    int kind;       /* KUN_BUF_* */
    int unreliable; /* warm-up rows of an output */
    int window;     /* rows of history consumers need */
} KunBufferDesc;

/* One kernel of a batch time-series stage: `ncomp` independent op groups
 * (components) in one flat grid. PAR kernels split each component's rows
 * into chunks (each chunk recomputes `comp_halo` rows of history before its
 * first stored row); SEQ kernels walk every row in order. The runtime sizes
 * chunks from `comp_cost` (estimated instructions per row) and passes the
 * launch table in kun::StageArgs. */
typedef struct {
    const void *kernel; /* host stub */
    int kind;           /* KUN_KERNEL_* */
    int ncomp;
    int halo;           /* largest comp_halo */
    int lds_bytes;      /* static LDS per workgroup */
    int io_values;      /* distinct HBM values read or written per (stock, row) */
    const int *comp_cost;
    const int *comp_halo;
    const int *comp_spill; /* HBM scratch bytes per (stock, chunk) column */
    const int *reads;      /* module buffers the kernel reads */
    int nreads;
    const int *writes;     /* module buffers the kernel writes */
    int nwrites;
    const int *kdeps;      /* kernels of the same stage whose materialised values it reads */
    int nkdeps;
    const int *mat_reads;  /* materialised values of the stage it reads / writes; value m is */
    int nmat_reads;        /* passed as kun::StageArgs::bufs[nin + nout + m] */
    const int *mat_writes;
    int nmat_writes;
} KunKernelDesc;

typedef struct {
    const char *name;
    int kind;             /* KUN_STAGE_* */
    const KunKernelDesc *kerns; /* batch kernels of a TS stage, run in order */
    int nkern;
    int nmat;             /* [rows][stocks] values handed between the stage's kernels */
    const void *stream_kernel; /* one-tick kernel host stub (STREAM modules); KUN_STAGE_GENERIC:
                                  its row kernel (kun::CsArgs), batch and stream */
    int block;            /* threads per workgroup */
    int state_bytes;      /* per-stock HBM bytes of carried state (STREAM modules) */
    const int *in;
    int nin;
    const int *out;
    int nout;
    const int *deps;
    int ndeps;
} KunStageDesc;

/* Stream-state wire table (STREAM modules): the parts of the reference's
 * serialised stream state (Runtime.cpp:534-546) in its order, and where
 * this backend keeps each of them. */
typedef struct {
    int stage;   /* time-series stage owning the ring */
    int tslot;   /* first T slot of the ring in the stage's state */
    int slots;   /* ring length, a power of two; tick t at slot t & (slots - 1) */
} KunRingRef;

typedef struct {
    const char *name;        /* buffer name in the reference's buffer table */
    int window;              /* rows the reference keeps */
    int buf;                 /* module buffer holding the newest row; -1: a stage's windowed temporary */
    const KunRingRef *rings; /* stage rings of the same values (newest row included) */
    int nrings;
} KunWireBuf;

enum { KUN_WS_FASTSUM = 0, KUN_WS_EMA = 1, KUN_WS_LINREG = 2, KUN_WS_ACCUM = 3, KUN_WS_SKIPLIST = 4 };

/* One state buffer of the reference (one object per blocking group). */
typedef struct {
    int kind;    /* KUN_WS_* */
    int stage;
    int tslot;   /* first T slot of the op's state */
    int islot;   /* first int32 slot of the op's state */
    int window;  /* window of the op (0 if none) */
    int tslots;  /* T slots of the stage's state; its int32 slots follow them */
} KunWireState;

typedef struct {
    unsigned abi;
    const char *name;
    int dtype;      /* KUN_DTYPE_* */
    int in_layout;  /* KUN_LAYOUT_* */
    int out_layout;
    int blocking;   /* STs blocking length */
    int aligned;    /* stock count must be a multiple of blocking */
    int nbuf;
    const KunBufferDesc *bufs;
    int nstage;
    const KunStageDesc *stages;
    const char *program; /* KUNPROG text of the module */
    const void *const *stream_inits; /* per stage: state-initialising kernel (STREAM modules) */
    /* STREAM modules, session pools: phases in launch order, and the kernel
     * initialising every stage's state over the lanes of a block table */
    int nphase;
    const KunPhaseDesc *phases;
    const void *pool_init;
    int nwire_buf;
    const KunWireBuf *wire_bufs;
    int nwire_state;
    const KunWireState *wire_states;
    const int *stage_tslots; /* STREAM modules: T slots of each stage's state (its int32 slots follow) */
} KunModuleDesc;

#endif
