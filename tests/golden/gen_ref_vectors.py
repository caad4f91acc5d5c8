"""Writes golden vectors from the reference's own CPU path.

Run in the build container only (needs /root/reference and g++):
    bash oracle/build_ref.sh && python tests/golden/gen_ref_vectors.py

For each case the reference's Python code generator (KunQuant.Driver) emits
C++ for a graph, g++ compiles it against the reference runtime built by
oracle/build_ref.sh (outputs stay in oracle/_ref/), and the module runs
through the reference C API (kunRunGraph / kunStream*) on seeded inputs.
Only inputs and outputs are committed, as .npz fixtures next to this script.

These vectors pin what no closed form pins: fp32 rounding of carried state
(Kahan rolling sums, EMA, regression sums), LogFast/Exp bit patterns, and
end-to-end factor values of the seeded random graphs."""
import ctypes
import os
import subprocess
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
REF = os.environ.get("KUN_REFERENCE", "/root/reference")
REF_OUT = os.path.join(ROOT, "oracle", "_ref")
sys.path.insert(0, ROOT)

from kunquant_amd.workloads import data as wdata  # noqa: E402
from kunquant_amd.workloads import graphs as wgraphs  # noqa: E402


def ref_namespace():
    sys.path.insert(0, REF)
    import KunQuant.Op as Op
    import KunQuant.Stage as Stage
    import KunQuant.ops as ops
    return wgraphs.namespace(Op, ops, Stage)


def ref_compile(K, f, name, cfg):
    from KunQuant.Driver import compileit
    srcs = compileit(f, name, **cfg)
    gen = os.path.join(REF_OUT, "gen")
    os.makedirs(gen, exist_ok=True)
    cpp = os.path.join(gen, f"{name}.cpp")
    with open(cpp, "w") as fh:
        fh.write(srcs[0])
    so = os.path.join(REF_OUT, f"{name}.so")
    subprocess.check_call(["g++", "-std=c++11", "-O3", "-mavx2", "-mfma", "-fPIC", "-shared",
                           "-fvisibility=hidden"] + CONTRACT + [ "-I", os.path.join(REF, "cpp"), cpp, "-o", so,
                           "-L", REF_OUT, "-lKunRuntime", f"-Wl,-rpath,{REF_OUT}"])
    return so


# The reference's JIT passes CppCompilerConfig.other_flags to g++; the golden
# vectors use ("-ffp-contract=off",) so that every a*b+c is rounded as two
# IEEE operations. With GCC's default for C++ (contraction on) results depend
# on which products GCC chooses to fuse.
CONTRACT = ["-ffp-contract=off"]

_capi = None


def capi():
    global _capi
    if _capi is None:
        L = ctypes.CDLL(os.path.join(REF_OUT, "libKunRuntime.so"), mode=ctypes.RTLD_GLOBAL)
        vp, sz, cp = ctypes.c_void_p, ctypes.c_size_t, ctypes.c_char_p
        for n, res, args in [
                ("kunCreateSingleThreadExecutor", vp, []), ("kunLoadLibrary", vp, [cp]),
                ("kunGetModuleFromLibrary", vp, [vp, cp]), ("kunCreateBufferNameMap", vp, []),
                ("kunSetBufferNameMap", None, [vp, cp, vp]), ("kunDestoryBufferNameMap", None, [vp]),
                ("kunRunGraph", None, [vp, vp, vp, sz, sz, sz, sz]), ("kunCreateStream", vp, [vp, vp, sz]),
                ("kunQueryBufferHandle", sz, [vp, cp]), ("kunStreamGetCurrentBuffer", vp, [vp, sz]),
                ("kunStreamPushData", None, [vp, sz, vp]), ("kunStreamRun", None, [vp]),
                ("kunDestoryStream", None, [vp])]:
            fn = getattr(L, n)
            fn.restype = res
            fn.argtypes = args
        _capi = L
    return _capi


def run_batch(so, name, inputs, outputs, S, T, start, L, dtype):
    C = capi()
    ex = C.kunCreateSingleThreadExecutor()
    lib = C.kunLoadLibrary(so.encode())
    mod = C.kunGetModuleFromLibrary(lib, name.encode())
    assert mod, name
    m = C.kunCreateBufferNameMap()
    res = {}
    for k, v in inputs.items():
        C.kunSetBufferNameMap(m, k.encode(), v.ctypes.data)
    for k in outputs:
        res[k] = np.full((L, S), np.nan, dtype=dtype)
        C.kunSetBufferNameMap(m, k.encode(), res[k].ctypes.data)
    C.kunRunGraph(ex, mod, m, S, T, start, L)
    C.kunDestoryBufferNameMap(m)
    return res


# The reference C API streams float data only; this harness drives the C++
# StreamContext (RunGraph.hpp:31-53) for both element types, including its
# state serialisation (Runtime.cpp:534-546) and resume (Runtime.cpp:424-467).
_HARNESS = r"""
#include <Kun/RunGraph.hpp>
#include <Kun/Module.hpp>
#include <Kun/IO.hpp>
#include <string.h>
using namespace kun;
extern "C" {
void *h_create(const void *mod, size_t S) {
    return new StreamContext(createSingleThreadExecutor(), (const Module *)mod, S);
}
void *h_create_from(const void *mod, size_t S, const char *blob, size_t n) {
    MemoryInputStream in(blob, n);
    return new StreamContext(createSingleThreadExecutor(), (const Module *)mod, S, &in);
}
size_t h_serialize(void *c, char *out, size_t cap) {
    MemoryOutputStream s;
    if (!((StreamContext *)c)->serializeStates(&s)) return (size_t)-1;
    if (out && s.getSize() <= cap) memcpy(out, s.getData(), s.getSize());
    return s.getSize();
}
size_t h_query(void *c, const char *n) { return ((StreamContext *)c)->queryBufferHandle(n); }
void h_push_f(void *c, size_t h, const float *d) { ((StreamContext *)c)->pushData(h, d); }
void h_push_d(void *c, size_t h, const double *d) { ((StreamContext *)c)->pushData(h, d); }
void h_run(void *c) { ((StreamContext *)c)->run(); }
const float *h_get_f(void *c, size_t h) { return ((StreamContext *)c)->getCurrentBufferPtrFloat(h); }
const double *h_get_d(void *c, size_t h) { return ((StreamContext *)c)->getCurrentBufferPtrDouble(h); }
void h_free(void *c) { delete (StreamContext *)c; }
}
"""


def harness():
    src = os.path.join(REF_OUT, "gen", "stream_harness.cpp")
    so = os.path.join(REF_OUT, "stream_harness.so")
    os.makedirs(os.path.dirname(src), exist_ok=True)
    with open(src, "w") as fh:
        fh.write(_HARNESS)
    subprocess.check_call(["g++", "-std=c++11", "-O2", "-mavx2", "-mfma", "-ffp-contract=off", "-fPIC", "-shared", "-I",
                           os.path.join(REF, "cpp"), src, "-o", so, "-L", REF_OUT, "-lKunRuntime",
                           f"-Wl,-rpath,{REF_OUT}"])
    H = ctypes.CDLL(so)
    vp, sz = ctypes.c_void_p, ctypes.c_size_t
    for n, res, args in [("h_create", vp, [vp, sz]), ("h_query", sz, [vp, ctypes.c_char_p]),
                         ("h_push_f", None, [vp, sz, vp]), ("h_push_d", None, [vp, sz, vp]), ("h_run", None, [vp]),
                         ("h_get_f", vp, [vp, sz]), ("h_get_d", vp, [vp, sz]), ("h_free", None, [vp]),
                         ("h_create_from", vp, [vp, sz, ctypes.c_char_p, sz]),
                         ("h_serialize", sz, [vp, ctypes.c_char_p, sz])]:
        fn = getattr(H, n)
        fn.restype = res
        fn.argtypes = args
    return H


def _serialize(H, st):
    n = H.h_serialize(st, None, 0)
    buf = ctypes.create_string_buffer(n)
    assert H.h_serialize(st, buf, n) == n
    return np.frombuffer(buf.raw, dtype=np.uint8).copy()


def run_stream(so, name, rows, outputs, S, dtype, snapshots=(), resume=None):
    """rows: {input name: [T, S]}. Returns ({output: [T, S]}, {tick: state
    blob after that many ticks}); with resume=(blob, t0) the context starts
    from the blob and is fed rows t0.. (outputs then hold those ticks)."""
    C = capi()
    H = harness()
    lib = C.kunLoadLibrary(so.encode())
    mod = C.kunGetModuleFromLibrary(lib, name.encode())
    t0 = 0
    if resume is not None:
        blob, t0 = resume
        st = H.h_create_from(mod, S, blob.tobytes(), blob.size)
    else:
        st = H.h_create(mod, S)
    hin = {k: H.h_query(st, k.encode()) for k in rows}
    hs = {k: H.h_query(st, k.encode()) for k in outputs}
    dbl = dtype == np.float64
    ct = ctypes.c_double if dbl else ctypes.c_float
    T = len(next(iter(rows.values())))
    res = {k: np.empty((T - t0, S), dtype=dtype) for k in outputs}
    blobs = {}
    for i in range(t0, T):
        for k, v in rows.items():
            r = np.ascontiguousarray(v[i], dtype=dtype)
            (H.h_push_d if dbl else H.h_push_f)(st, hin[k], r.ctypes.data)
        H.h_run(st)
        for k in outputs:
            p = (H.h_get_d if dbl else H.h_get_f)(st, hs[k])
            res[k][i - t0] = np.ctypeslib.as_array(ctypes.cast(p, ctypes.POINTER(ct)), shape=(S,))
        if i + 1 in snapshots:
            blobs[i + 1] = _serialize(H, st)
    H.h_free(st)
    return res, blobs


def output_names(f):
    return [op.attrs["name"] for op in f.ops if op.__class__.__name__ == "Output"]


def ops_inputs(S, T, dtype, seed):
    rng = np.random.Generator(np.random.PCG64(seed))
    a = (rng.random((T, S)) + 0.5).astype(dtype)
    b = (rng.random((T, S)) - 0.3).astype(dtype)
    a[7, 3] = np.nan
    b[11, 5] = np.nan
    a[20:23, :] = a[19, :]  # ties in windows
    return {"a": a, "b": b}


def alpha158_ref_graph(K):
    sys.path.insert(0, REF)
    from KunQuant.predefined.Alpha158 import AllData
    return wgraphs.alpha158_graph(K, AllData)


# stocks whose columns the large-universe chain fixture keeps
CHAIN_STOCKS = [128, 4774, 4775]


def alpha101_ref_graph(K):
    sys.path.insert(0, REF)
    import KunQuant.predefined.Alpha101 as lib
    return wgraphs.alpha101_graph(K, library=lib)


def alpha101_inputs(dtype, seed):
    """Prices and volumes with the share of flat moves real data has (equal
    closes, high == low days) so that the 0.0001 divisor guards and ties
    are exercised."""
    d = wdata.ohlcv(20, 320, seed=seed, dtype=dtype, nan_frac=0.0)
    rng = np.random.Generator(np.random.PCG64(seed + 1))
    flat = rng.random((320, 20)) < 0.03
    for k in ("open", "high", "low"):
        d[k] = np.where(flat, d["close"], d[k]).astype(dtype)
    return d


def large_window_inputs(S, T, seed):
    """Uniform values with a block of repeated minima and NaN cells (ties
    and gaps inside 200-row windows)."""
    rng = np.random.Generator(np.random.PCG64(seed))
    a = rng.random((T, S))
    a[400:410, :] = -1
    a[rng.random((T, S)) < 0.002] = np.nan
    return {"a": a}


def generic_inputs(S, T, dtype, seed):
    rng = np.random.Generator(np.random.PCG64(seed))
    d = {k: rng.random((T, S)).astype(dtype) for k in ("a", "b", "c")}
    d["a"][5, 7] = np.nan
    return d


CASES = [
    # name, graph factory, compile config, dtype, inputs(S, T), S, T, start, L
    # the reference's own Alpha158 library, compiled as its test_alpha158.py:40-43 does
    ("alpha158_f64", alpha158_ref_graph,
     dict(dtype="double", blocking_len=4, partition_factor=4, input_layout="TS", output_layout="TS",
          options={"opt_reduce": True, "fast_log": True}),
     np.float64, lambda: wdata.ohlcv(16, 200, seed=158, dtype=np.float64, nan_frac=0.02), 16, 200, 0, 200),
    # the reference's own Alpha101 library, compiled as its test_alpha101.py:27-47 does
    ("alpha101_f32", alpha101_ref_graph, dict(input_layout="TS", output_layout="TS",
                                              options={"opt_reduce": True, "fast_log": True}),
     np.float32, lambda: alpha101_inputs(np.float32, 101), 20, 320, 0, 320, 256),
    ("alpha101_f64", alpha101_ref_graph, dict(dtype="double", input_layout="TS", output_layout="TS",
                                              options={"opt_reduce": True, "fast_log": True,
                                                       "no_fast_stat": "no_warn"}),
     np.float64, lambda: alpha101_inputs(np.float64, 102), 20, 320, 0, 320, 256),
    # 200-row windows: the reference's skip-list forms (fp64, test_runtime.py:271-300's shape)
    ("skiplist_f64", wgraphs.large_window_graph, dict(dtype="double", input_layout="TS", output_layout="TS",
                                                      options={"no_fast_stat": "no_warn"}),
     np.float64, lambda: large_window_inputs(24, 1200, 44), 24, 1200, 0, 1200, 900),
    ("generic_f32", wgraphs.generic_cs_graph, dict(input_layout="TS", output_layout="TS"), np.float32,
     lambda: generic_inputs(40, 60, np.float32, 21), 40, 60, 0, 60),
    ("simple_f32", wgraphs.simple_graph, dict(input_layout="TS", output_layout="TS"), np.float32,
     lambda: {k: v for k, v in wdata.ohlcv(64, 260, seed=1).items() if k != "amount"}, 64, 260, 0, 260),
    ("ops_f32", wgraphs.ops_graph, dict(input_layout="TS", output_layout="TS"), np.float32,
     lambda: ops_inputs(40, 120, np.float32, 7), 40, 120, 10, 110),
    ("ops_f64", wgraphs.ops_graph, dict(dtype="double", input_layout="TS", output_layout="TS",
                                        options={"no_fast_stat": "no_warn"}), np.float64,
     lambda: ops_inputs(40, 120, np.float64, 8), 40, 120, 10, 110),
    ("rand_f32", lambda K: wgraphs.random_graph(100, 102, K=K), dict(input_layout="TS", output_layout="TS"),
     np.float32, lambda: wdata.ohlcv(24, 130, seed=2, nan_frac=0.1), 24, 130, 0, 130),
    ("rand_f64", lambda K: wgraphs.random_graph(40, 103, K=K),
     dict(dtype="double", input_layout="TS", output_layout="TS", options={"no_fast_stat": "no_warn"}),
     np.float64, lambda: wdata.ohlcv(24, 130, seed=3, dtype=np.float64, nan_frac=0.1), 24, 130, 0, 130),
]

STREAM_CASES = [
    ("stream_f64", dict(dtype="double", input_layout="STREAM", output_layout="STREAM"), np.float64),
    ("stream_f32", dict(input_layout="STREAM", output_layout="STREAM", options={"opt_reduce": False}), np.float32),
]


AGG_KINDS = ("sum", "min", "max", "first", "last", "count", "mean")


def rowreduce_vectors():
    """corrWith (TS and STs, plain and ranked inputs) and aggregrate (float,
    double) through the reference runtime's own functions (a harness built
    from tests/golden/ref_rowreduce.cpp): tests/golden/ref_rowreduce.npz."""
    so = os.path.join(REF_OUT, "ref_rowreduce.so")
    subprocess.check_call(["g++", "-std=c++11", "-O2", "-fPIC", "-shared"] + CONTRACT +
                          ["-I", os.path.join(REF, "cpp"), os.path.join(HERE, "ref_rowreduce.cpp"), "-o", so,
                           "-L", REF_OUT, "-lKunRuntime", f"-Wl,-rpath,{REF_OUT}"])
    capi()  # loads libKunRuntime.so globally
    H = ctypes.CDLL(so)
    fp = ctypes.POINTER(ctypes.c_void_p)
    vp, sz, i = ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int
    H.ref_corr_with.argtypes = [i, i, i, fp, vp, fp, sz, sz, sz, sz]
    H.ref_aggregrate.argtypes = [i, fp, fp, i, fp, sz, sz, sz, sz]
    ptrs = lambda arrs: (vp * len(arrs))(*[None if a is None else a.ctypes.data for a in arrs])
    rng = np.random.Generator(np.random.PCG64(77))
    out = {}
    # corrWith: S = 37 (not a multiple of 8 in TS), 60 rows, one all-NaN row
    # and NaN cells (plain: NaN rows out; ranked: NaN cells drop out of the
    # ranks), ties for the ranked path
    S, T = 37, 60
    a = rng.random((T, S)).astype(np.float32)
    b = np.floor(rng.random((T, S)) * 6).astype(np.float32)
    r = rng.random((T, S)).astype(np.float32)
    a[7, 3] = np.nan
    b[11, :] = np.nan
    out.update(cw_a=a, cw_b=b, cw_r=r)
    for rank in (0, 1):
        o = [np.full(T, -7, np.float32) for _ in range(2)]
        H.ref_corr_with(0, rank, 2, ptrs([a, b]), r.ctypes.data, ptrs(o), S, T, 0, T)
        out[f"cw_ts_rank{rank}"] = np.stack(o)
    # STs layout: 40 stocks (blocks of 8), rows 5..59
    S2 = 40
    a2 = rng.random((T, S2)).astype(np.float32)
    r2 = rng.random((T, S2)).astype(np.float32)
    sts = lambda x: np.ascontiguousarray(x.reshape(T, S2 // 8, 8).transpose(1, 0, 2))
    o = [np.full(T - 5, -7, np.float32)]
    a2s, r2s = sts(a2), sts(r2)   # kept alive across the call
    H.ref_corr_with(1, 0, 1, ptrs([a2s]), r2s.ctypes.data, ptrs(o), S2, T, 5, T - 5)
    out.update(cw_sts_a=a2, cw_sts_r=r2, cw_sts=o[0])
    # aggregrate: 19 stocks, 200 rows in 9 label groups of uneven length,
    # NaN cells inside groups
    for dt, name in ((np.float32, "f32"), (np.float64, "f64")):
        S3, T3 = 19, 200
        x = (rng.standard_normal((T3, S3)) * 10).astype(dt)
        x[17, 3] = np.nan
        x[101:104, 5] = np.nan
        cuts = np.sort(rng.choice(np.arange(1, T3), 8, replace=False))
        labels = np.zeros(T3, dt)
        for c in cuts:
            labels[c:] += 1
        groups = 9
        outs = [np.full((groups, S3), -7, dt) for _ in AGG_KINDS]
        H.ref_aggregrate(1, ptrs([x]), ptrs([labels]), 1 if dt == np.float64 else 0, ptrs(outs), S3, T3, 0, T3)
        out[f"agg_{name}_x"] = x
        out[f"agg_{name}_labels"] = labels
        for k, v in zip(AGG_KINDS, outs):
            out[f"agg_{name}_{k}"] = v
    np.savez_compressed(os.path.join(HERE, "ref_rowreduce.npz"), **out)
    print("wrote rowreduce", len(out), "arrays")


def main():
    if not os.path.exists(os.path.join(REF_OUT, "libKunRuntime.so")):
        subprocess.check_call(["bash", os.path.join(ROOT, "oracle", "build_ref.sh")])
    if "--rowreduce" in sys.argv:
        rowreduce_vectors()
        return
    K = ref_namespace()
    if "--bench-modules" in sys.argv:
        # reference builds of bench graphs with no fixture: bench.py times
        # them as its CPU baseline and compares their outputs with the
        # device's (oracle/_ref only, nothing committed)
        by_case = {c[0]: c for c in CASES}
        for name, factory, cfg in [
                ("rand_f32", by_case["rand_f32"][1], by_case["rand_f32"][2]),
                ("alpha158_f64", by_case["alpha158_f64"][1], by_case["alpha158_f64"][2]),
                ("cfg2_f64", lambda K: wgraphs.random_graph(100, 102, K=K),
                 dict(dtype="double", input_layout="TS", output_layout="TS", options={"no_fast_stat": "no_warn"})),
                ("prec_f32", lambda K: wgraphs.precision_graph(150, 3, K=K),
                 dict(input_layout="TS", output_layout="TS")),
                ("prec_f64", lambda K: wgraphs.precision_graph(150, 3, K=K),
                 dict(dtype="double", input_layout="TS", output_layout="TS", options={"no_fast_stat": True}))]:
            ref_compile(K, factory(K), name, dict(cfg))
            print("built reference module", name)
        return
    # --stream-only: rewrite the stream fixtures alone; --only=<case>: one batch case
    only = [a.split("=", 1)[1] for a in sys.argv if a.startswith("--only=")]
    batch_cases = [] if "--stream-only" in sys.argv else [c for c in CASES if not only or c[0] in only]
    for name, factory, cfg, dtype, make_inputs, S, T, start, L, *keep in batch_cases:
        f = factory(K)
        outs = output_names(f)
        so = ref_compile(K, f, name, dict(cfg))
        inputs = make_inputs()
        res = run_batch(so, name, inputs, outs, S, T, start, L, dtype)
        # large libraries: only output rows keep[0].. are stored (meta[4])
        k0 = keep[0] if keep else 0
        np.savez_compressed(os.path.join(HERE, f"ref_{name}.npz"),
                            **{f"in_{k}": v for k, v in inputs.items()},
                            **{f"out_{k}": np.ascontiguousarray(v[k0:]) for k, v in res.items()},
                            meta=np.array([S, T, start, L] + ([k0] if keep else [])))
        print("wrote", name, len(outs), "outputs")
    if only and "tsrank_chain" in only:
        # the bench's volume (8192 stocks x 1024 days, regenerated from its
        # seed by the test); only three stocks' columns are stored
        f = wgraphs.tsrank_chain_graph(K)
        outs = output_names(f)
        so = ref_compile(K, f, "tsrank_chain", dict(input_layout="TS", output_layout="TS"))
        vol = wdata.ohlcv(8192, 1024, seed=2, dtype=np.float32, nan_frac=0.1)["volume"]
        res = run_batch(so, "tsrank_chain", {"volume": np.ascontiguousarray(vol)}, outs, 8192, 1024, 0, 1024,
                        np.float32)
        np.savez_compressed(os.path.join(HERE, "ref_tsrank_chain_f32.npz"),
                            **{f"out_{k}": v[:, CHAIN_STOCKS] for k, v in res.items()},
                            stocks=np.array(CHAIN_STOCKS), meta=np.array([8192, 1024, 2]))
        print("wrote tsrank_chain")
    if only:
        return
    rng = np.random.Generator(np.random.PCG64(11))
    for name, cfg, dtype in STREAM_CASES:
        f = wgraphs.stream_graph(K)
        outs = output_names(f)
        so = ref_compile(K, f, name, dict(cfg))
        S = 16
        rows = rng.random((60, S)).astype(dtype)
        rows[13, 2] = np.nan
        res, blobs = run_stream(so, name, {"a": rows}, outs, S, dtype, snapshots=(37, 60))
        # resuming from the blob reproduces the uninterrupted run
        tail, _ = run_stream(so, name, {"a": rows}, outs, S, dtype, resume=(blobs[37], 37))
        for k in outs:
            assert np.array_equal(tail[k], res[k][37:], equal_nan=True), k
        np.savez_compressed(os.path.join(HERE, f"ref_{name}.npz"), in_a=rows,
                            **{f"out_{k}": v for k, v in res.items()}, meta=np.array([S, len(rows)]),
                            state37=blobs[37], state60=blobs[60])
        print("wrote", name, "state blobs", blobs[37].size, blobs[60].size)
    # stream-state blobs of the 100-factor serving graph (windowed stage
    # temporaries, rolling sums, regressions and skip lists across stages)
    name, S, T, K1 = "wire_serve_f32", 16, 90, 71
    f = wgraphs.random_graph(100, 5000, K=K)
    outs = output_names(f)
    so = ref_compile(K, f, name, dict(input_layout="STREAM", output_layout="STREAM", options={"opt_reduce": False}))
    d = wdata.ohlcv(S, T, seed=12, dtype=np.float32, nan_frac=0.05)
    ins = [op.attrs["name"] for op in f.ops if op.__class__.__name__ == "Input"]
    rows = {k: d[k] for k in ins}
    res, blobs = run_stream(so, name, rows, outs, S, np.float32, snapshots=(K1, T))
    np.savez_compressed(os.path.join(HERE, f"ref_{name}.npz"), **{f"in_{k}": v for k, v in rows.items()},
                        **{f"out_{k}": v[-3:] for k, v in res.items()}, meta=np.array([S, T, K1]),
                        state_k1=blobs[K1], state_end=blobs[T])
    print("wrote", name, "state blobs", blobs[K1].size, blobs[T].size)


if __name__ == "__main__":
    main()
