"""Device results against the reference's golden vectors and the oracle.

Bar: every output value equal (bit-exact for ranks, arg-indices, masks and
all floating-point values; NaN positions equal, NaN payload bits ignored).
Runs through the C ABI (libkunhip.so) via KunRunner."""
import os

import numpy as np
import pytest

from kunquant_amd.codegen.program import program_text
from kunquant_amd.runner import KunRunner as kr
from kunquant_amd.workloads import catalog, data
from oracle.kun_oracle import OracleProgram

pytestmark = pytest.mark.gpu

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
NTHREADS = 16


@pytest.fixture(scope="module")
def executor():
    return kr.createSingleThreadExecutor()


def golden(name):
    z = np.load(os.path.join(GOLDEN, f"ref_{name}.npz"))
    return ({k[3:]: z[k] for k in z.files if k.startswith("in_")},
            {k[4:]: z[k] for k in z.files if k.startswith("out_")}, [int(x) for x in z["meta"]])


def assert_same(got, exp, what):
    bad = [k for k in exp if not np.array_equal(got[k], exp[k], equal_nan=True)]
    if bad:
        k = bad[0]
        a, b = got[k], exp[k]
        fin = ~np.isnan(a) & ~np.isnan(b)
        msg = (f"{what}: {len(bad)}/{len(exp)} outputs differ (first: {k}, NaN masks equal: "
               f"{np.array_equal(np.isnan(a), np.isnan(b))}, values differing: {int((a[fin] != b[fin]).sum())})")
        raise AssertionError(msg)


@pytest.mark.parametrize("golden_name,module", [("simple_f32", "simple_f32"), ("ops_f32", "ops_f32"),
                                                ("ops_f64", "ops_f64"), ("rand_f32", "cfg2_f32"),
                                                ("rand_f64", "rand40_f64"),
                                                # 200-row windows: the skip-list forms of TsRank,
                                                # Min / Max and ArgMin / ArgMax (fp64)
                                                ("skiplist_f64", "skiplist_f64")])
def test_device_matches_reference_golden(executor, golden_name, module):
    ins, outs, meta = golden(golden_name)
    S, T, start, L = meta[:4]
    k0 = meta[4] if len(meta) > 4 else 0   # first stored output row
    _, mod = catalog.load(module)
    got = kr.runGraph(executor, mod, {k: ins[k] for k in mod.getInputNames()}, start, L)
    assert_same({k: v[k0:] for k, v in got.items()}, outs, f"{module} vs reference")


def _oracle(module):
    return OracleProgram(program_text(catalog.module_ir(module)))


@pytest.mark.parametrize("module,S,T,start,L,nan_frac", [
    ("cfg2_f32", 1000, 300, 17, 250, 0.1),   # unaligned stocks, cur_time > 0, NaN cells
    ("cfg2_f64", 333, 200, 0, 200, 0.1),     # fp64 fast statistics
    ("simple_f32", 8, 40, 0, 40, 0.0),       # fewer stocks than a wavefront
    ("simple_f32", 513, 70, 5, 60, 0.02),
])
def test_device_matches_oracle(executor, module, S, T, start, L, nan_frac):
    _, mod = catalog.load(module)
    d = data.ohlcv(S, T, seed=S + T, dtype=mod.dtype, nan_frac=nan_frac)
    ins = {k: d[k] for k in mod.getInputNames()}
    got = kr.runGraph(executor, mod, ins, start, L)
    exp = _oracle(module).run(ins, S, T, start, L, nthreads=NTHREADS)
    assert_same(got, exp, f"{module} {S}x{T}")


@pytest.mark.parametrize("module", ["ops_f32", "ops_f64"])
def test_ops_edge_values(executor, module):
    """Ties, NaN cells, zeros and infinities through every primitive."""
    _, mod = catalog.load(module)
    S, T = 70, 90
    rng = np.random.default_rng(21)
    a = np.round(rng.random((T, S)) * 4, 1).astype(mod.dtype) + 0.5
    b = (rng.random((T, S)) - 0.5).astype(mod.dtype)
    a[rng.random((T, S)) < 0.03] = np.nan
    b[rng.random((T, S)) < 0.03] = np.nan
    b[5, :] = 0.0
    a[9, :3] = np.inf
    b[12, 4:8] = -0.0
    ins = {"a": a, "b": b}
    got = kr.runGraph(executor, mod, ins, 0, T)
    exp = _oracle(module).run(ins, S, T, 0, T, nthreads=NTHREADS)
    assert_same(got, exp, module)


def test_sts_layout(executor):
    """STs input/output layout ([S/8, T, 8]); padding lanes are not compared."""
    _, mod = catalog.load("ops_f32_sts")
    S, T, b = 100, 80, mod.blocking_len
    Sp = (S + b - 1) // b * b
    rng = np.random.default_rng(22)
    ts = {"a": rng.random((T, S)).astype(np.float32) + 0.5, "b": rng.random((T, S)).astype(np.float32) - 0.3}

    def to_sts(x):
        xp = np.zeros((T, Sp), dtype=x.dtype)
        xp[:, :S] = x
        return np.ascontiguousarray(xp.reshape(T, Sp // b, b).transpose(1, 0, 2))

    ins = {k: to_sts(v) for k, v in ts.items()}
    got = kr.runGraph(executor, mod, ins, 0, T, num_stocks=S)
    exp = _oracle("ops_f32_sts").run(ins, S, T, 0, T, nthreads=NTHREADS)
    unb = lambda x: x.transpose(1, 0, 2).reshape(T, Sp)[:, :S]
    assert_same({k: unb(v) for k, v in got.items()}, {k: unb(v) for k, v in exp.items()}, "STs")


@pytest.mark.parametrize("module,policy", [
    ("cfg2_f32", (10 ** 6, 0, 1)),   # one row per chunk: every row rebuilds its whole halo
    ("cfg2_f32", (4096, 1, 7)),      # ragged 7-row chunks
    ("ops_f32", (10 ** 6, 0, 1)),
    ("ops_f64", (4096, 1, 5)),
])
def test_row_chunking_matches_oracle(module, policy):
    """Row-chunked kernels give the same bits however the rows are split."""
    ex = kr.createSingleThreadExecutor()
    ex.set_chunking(*policy)
    _, mod = catalog.load(module)
    S, T, start, L = 130, 160, 9, 150
    d = data.ohlcv(S, T, seed=5, dtype=mod.dtype, nan_frac=0.05)
    cols = list(d.values())
    ins = {k: d.get(k, cols[i % len(cols)]) for i, k in enumerate(mod.getInputNames())}
    got = kr.runGraph(ex, mod, ins, start, L)
    exp = _oracle(module).run(ins, S, T, start, L, nthreads=NTHREADS)
    assert_same(got, exp, f"{module} chunking {policy}")


def test_full_size_config2(executor):
    """8192 stocks x 1024 rows against the oracle over every row; a second
    run is bit-identical."""
    _, mod = catalog.load("cfg2_f32")
    S, T = 8192, 1024
    d = data.ohlcv(S, T, seed=2, nan_frac=0.1)
    got = kr.runGraph(executor, mod, d, 0, T)
    again = kr.runGraph(executor, mod, d, 0, T)
    for k in got:
        assert np.array_equal(got[k].view(np.uint32), again[k].view(np.uint32)), k
    exp = _oracle("cfg2_f32").run(d, S, T, 0, T, nthreads=NTHREADS)
    assert_same(got, exp, "config-2 full size")


def _cs_rows(S, dt, seed):
    rng = np.random.default_rng(seed)
    rows = [np.full(S, np.nan), np.zeros(S), np.round(rng.random(S) * 3), rng.choice([-0.0, 0.0, 1.0], S),
            rng.choice([-np.inf, np.inf, 2.0, np.nan], S), np.where(np.arange(S) == S // 2, 5.0, np.nan)]
    for _ in range(6):
        x = rng.standard_normal(S)
        x[rng.random(S) < 0.1] = np.nan
        rows.append(x)
    # magnitudes over many binades (every key byte differs across the row)
    # and near ties one unit in the last place apart (only the low bytes do)
    rows.append(rng.standard_normal(S) * 10.0 ** rng.uniform(-30, 30, S))
    one = np.asarray(1.0, dt)
    rows.append(one + rng.integers(0, 5, S) * np.finfo(dt).eps)
    # Scale's quotient exactly between two subnormal floats: the row sums to
    # c = 2C (the subnormal cells vanish in the float sum) and x / c = o *
    # 2^-150 for odd o, which rounds to the even neighbour
    for C in (8175, 16307):
        o = 2 * (((1 << 24) // C - 1) // 2) + 1 - 2 * np.arange(S - 1)
        o = o[o > 0]
        row = np.zeros(S)
        row[0] = 2.0 * C
        row[1:1 + len(o)] = (o * C).astype(np.float64) * 2.0 ** -149
        rows.append(row)
    return np.ascontiguousarray(np.stack(rows).astype(dt))


@pytest.mark.parametrize("module", ["cs_f32", "cs_f64"])
@pytest.mark.parametrize("S", [1, 5, 64, 257, 1000, 2000, 4000, 8192, 16384])
def test_cross_section_edges(executor, module, S):
    """Rank / Scale rows: all NaN, all zero, heavy ties, signed zeros,
    infinities, a single valid value, random rows with NaN cells."""
    _, mod = catalog.load(module)
    a = _cs_rows(S, mod.dtype, S)
    b = _cs_rows(S, mod.dtype, S + 1)[::-1].copy()
    ins = {"a": a, "b": b}
    T = a.shape[0]
    got = kr.runGraph(executor, mod, ins, 0, T)
    exp = _oracle(module).run(ins, S, T, 0, T, nthreads=NTHREADS)
    assert_same(got, exp, f"{module} S={S}")
