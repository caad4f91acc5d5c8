"""corrWith and aggregrate on the device (test_runtime.py:17-122): equal
to the oracle bit for bit (NaN positions equal) and to pandas at the
reference test's tolerance, through KunRunner (host buffers) and the
device-pointer entry points."""
import ctypes

import numpy as np
import pytest

from kunquant_amd.runner import KunRunner as kr
from oracle import kun_oracle as ko

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")


@pytest.fixture(scope="module")
def executor():
    return kr.createSingleThreadExecutor()


def _same(a, b):
    return np.array_equal(a, b, equal_nan=True)


@pytest.mark.parametrize("S,T", [(20, 24), (1000, 64), (8192, 40)])
@pytest.mark.parametrize("rank_inputs", [False, True])
def test_corrwith_matches_oracle(executor, S, T, rank_inputs):
    rng = np.random.default_rng(S + T)
    a = rng.random((T, S)).astype(np.float32)
    b = rng.random((T, S)).astype(np.float32)
    r = rng.random((T, S)).astype(np.float32)
    a[2, 5] = np.nan                       # a NaN row (plain) / NaN rank cell
    b[:, 3] = b[:, 4]                      # ties for the rank path
    o1, o2 = np.empty(T, np.float32), np.empty(T, np.float32)
    kr.corrWith(executor, [a, b], r, [o1, o2], "TS", rank_inputs=rank_inputs)
    assert _same(o1, ko.corr_with(a, r, rank_inputs=rank_inputs))
    assert _same(o2, ko.corr_with(b, r, rank_inputs=rank_inputs))


def test_corrwith_sts_layout(executor):
    rng = np.random.default_rng(5)
    S, T = 64, 30
    a = rng.random((T, S)).astype(np.float32)
    r = rng.random((T, S)).astype(np.float32)
    sts = lambda x: np.ascontiguousarray(x.reshape(T, S // 8, 8).transpose(1, 0, 2))
    o = np.empty(T, np.float32)
    kr.corrWith(executor, [sts(a)], sts(r), [o], "STs")
    assert _same(o, ko.corr_with(a, r))


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_aggregrate_matches_oracle_and_pandas(executor, dtype):
    import pandas as pd
    rng = np.random.default_rng(7)
    a = rng.random((240, 16)).astype(dtype)
    a[17, 3] = np.nan
    dates = pd.date_range(start="2026-01-01", end="2026-01-07", freq="min")
    stamps = pd.DatetimeIndex(sorted(rng.choice(dates, size=240)))
    labels = stamps.day.to_numpy().astype(dtype)
    groups = len(np.unique(stamps.date))
    outs = {k: np.empty((groups, 16), dtype) for k in kr.AGGREGRATION_KINDS}
    kr.aggregrate(executor, [a], [labels], [outs])
    exp = ko.aggregate(a, labels)
    for k in kr.AGGREGRATION_KINDS:
        assert _same(outs[k], exp[k]), k
    clean = a.copy()
    clean[17, 3] = 0.5
    outs2 = {"sum": np.empty((groups, 16), dtype), "mean": np.empty((groups, 16), dtype)}
    kr.aggregrate(executor, [clean], [labels], [outs2])
    g = pd.DataFrame(clean, index=stamps).groupby(stamps.date)
    np.testing.assert_allclose(outs2["sum"], g.sum(), atol=1e-6, rtol=1e-4)
    np.testing.assert_allclose(outs2["mean"], g.mean(), atol=1e-6, rtol=1e-4)


def test_aggregrate_device_entry_checks_rows(executor):
    """The device entry point counts the label groups and refuses outputs
    with fewer rows."""
    L = kr._load()
    S, T = 64, 50
    x = torch.rand((T, S), device="cuda:0")
    lab = torch.repeat_interleave(torch.arange(5, device="cuda:0", dtype=torch.float32), 10)
    out = torch.empty((5, S), device="cuda:0")
    vp = ctypes.c_void_p
    outs = (vp * 7)(out.data_ptr(), *([None] * 6))
    rc = L.kunAggregrateDevice(executor._h, 1, (vp * 1)(x.data_ptr()), (vp * 1)(lab.data_ptr()), 0, outs, 5, S, T,
                               0, T, None)
    assert rc == 0, kr._err()
    torch.cuda.synchronize()
    exp = ko.aggregate(x.cpu().numpy(), lab.cpu().numpy())["sum"]
    assert _same(out.cpu().numpy(), exp)
    rc = L.kunAggregrateDevice(executor._h, 1, (vp * 1)(x.data_ptr()), (vp * 1)(lab.data_ptr()), 0, outs, 4, S, T,
                               0, T, None)
    assert rc != 0 and "exceed" in kr._err()


def _ref():
    import os
    return np.load(os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "ref_rowreduce.npz"))


@pytest.mark.parametrize("rank_inputs", [False, True])
def test_corrwith_matches_reference_vectors(executor, rank_inputs):
    """corrWith on the device against the reference runtime's own corrWith
    (tests/golden/ref_rowreduce.npz: TS rows of 37 stocks with NaN cells, a
    NaN row and ties; STs rows 5..59 of 40 stocks), bit for bit."""
    z = _ref()
    T = z["cw_a"].shape[0]
    o = [np.empty(T, np.float32) for _ in range(2)]
    kr.corrWith(executor, [z["cw_a"], z["cw_b"]], z["cw_r"], o, "TS", rank_inputs=rank_inputs)
    exp = z[f"cw_ts_rank{int(rank_inputs)}"]
    assert _same(o[0], exp[0]) and _same(o[1], exp[1])
    if not rank_inputs:
        a2, r2 = z["cw_sts_a"][5:], z["cw_sts_r"][5:]
        Ts, S2 = a2.shape
        sts = lambda x: np.ascontiguousarray(x.reshape(Ts, S2 // 8, 8).transpose(1, 0, 2))
        o2 = np.empty(Ts, np.float32)
        kr.corrWith(executor, [sts(a2)], sts(r2), [o2], "STs")
        assert _same(o2, z["cw_sts"])


@pytest.mark.parametrize("dt", ["f32", "f64"])
def test_aggregrate_matches_reference_vectors(executor, dt):
    """aggregrate on the device against the reference runtime's own
    aggregrate (19 stocks, 200 rows in 9 uneven label groups, NaN cells),
    every kind bit for bit."""
    z = _ref()
    x, labels = z[f"agg_{dt}_x"], z[f"agg_{dt}_labels"]
    outs = {k: np.empty(z[f"agg_{dt}_{k}"].shape, x.dtype) for k in kr.AGGREGRATION_KINDS}
    kr.aggregrate(executor, [x], [labels], [outs])
    for k in kr.AGGREGRATION_KINDS:
        assert _same(outs[k], z[f"agg_{dt}_{k}"]), k
