"""Pins the CPU oracle to closed forms and to the reference's own runtime
test cases (KunQuant tests/test_runtime.py, tests/test_stream.py).

Oracle functions and what pins them:
  rolling sum/mean/std/var/cov/corr/skew/kurt   pandas rolling (test_runtime.py:187-247, 614-630)
  TsRank/TsArgMin/TsArgMax/min/max              pandas rolling, with duplicates (test_runtime.py:271-300, 366-381)
  EMA (NaN row, nested, __init)                 pandas ewm (test_runtime.py:304-351)
  Rank (ties, NaN rows, bit-exact)              pandas rank pct (test_runtime.py:393-489)
  Scale                                          df.div(df.abs().sum(axis=1)) (ref_alpha101.py:147-154)
  quantile                                       pandas rolling quantile (test_runtime.py:259-268)
  regression slope / R^2 / residual             numpy polyfit / closed forms (test_stream.py:46)
  DecayLinear                                    weighted sum closed form (ref_alpha101.py:174-198)
  Log / Exp / Pow                                numpy at rtol 1e-5 (test_runtime.py:501-561)
  max drawdown                                   brute force (test_runtime.py:737-771)
  NaN warm-up masks                              unreliable counts (InferWindow.py:6-23)
A plausible slip (wrong window offset, dropped Kahan term, transposed
operand, wrong tie rule) fails at least one of these."""
import numpy as np
import pandas as pd
import pytest

from kunquant_amd.Op import Builder, ConstantOp, Input, Output, Rank, Scale
from kunquant_amd.Stage import Function
from kunquant_amd.ops import *
from tests.helpers import oracle_module, run_oracle


def graph(build):
    b = Builder()
    with b:
        build()
    return Function(b.ops)


def test_avg_stddev_sts_float():  # test_runtime.py:187-202 (STs, fp32, two-pass)
    f = graph(lambda: (Output(WindowedAvg(Input("a"), 10), "ou1"), Output(WindowedStddev(Input("a"), 10), "ou2")))
    inp = np.random.default_rng(0).random((24, 20)).astype("float32")   # [S, T]
    blocked = np.ascontiguousarray(inp.reshape((-1, 8, 20)).transpose((0, 2, 1)))
    m, p = oracle_module(f, input_layout="STs", output_layout="STs")
    out = p.run({"a": blocked}, 24, 20, 0, 20)
    unblock = lambda x: np.ascontiguousarray(x.transpose((0, 2, 1)).reshape((-1, 20)))
    df = pd.DataFrame(inp.transpose())
    np.testing.assert_allclose(unblock(out["ou1"]), df.rolling(10).mean().to_numpy().T, rtol=1e-6, equal_nan=True)
    np.testing.assert_allclose(unblock(out["ou2"]), df.rolling(10).std().to_numpy().T, rtol=1e-5, equal_nan=True)


@pytest.mark.parametrize("fast", [False, True])
def test_avg_stddev_ts_double(fast):  # test_runtime.py:206-225 (fast and two-pass statistics)
    f = graph(lambda: (Output(WindowedAvg(Input("a"), 10), "ou1"), Output(WindowedStddev(Input("a"), 10), "ou2"),
                       Output(WindowedVar(Input("a"), 10), "ou3")))
    inp = np.random.default_rng(1).random((200, 20))
    out = run_oracle(f, {"a": inp}, dtype="double", options={"no_fast_stat": "no_warn" if fast else True})
    df = pd.DataFrame(inp)
    np.testing.assert_allclose(out["ou1"], df.rolling(10).mean(), rtol=1e-9, equal_nan=True)
    np.testing.assert_allclose(out["ou2"], df.rolling(10).std(), rtol=1e-8, equal_nan=True)
    np.testing.assert_allclose(out["ou3"], df.rolling(10).var(), rtol=1e-8, equal_nan=True)


@pytest.mark.parametrize("fast", [False, True])
def test_cov_corr(fast):  # test_runtime.py:229-247
    f = graph(lambda: (Output(WindowedCovariance(Input("a"), 10, Input("b")), "ou1"),
                       Output(WindowedCorrelation(Input("a"), 10, Input("b")), "ou2")))
    rng = np.random.default_rng(2)
    a, b = rng.random((200, 24)), rng.random((200, 24))
    out = run_oracle(f, {"a": a, "b": b}, dtype="double", options={"no_fast_stat": "no_warn" if fast else True})
    da, db = pd.DataFrame(a), pd.DataFrame(b)
    np.testing.assert_allclose(out["ou1"], da.rolling(10).cov(db), rtol=1e-6, equal_nan=True)
    np.testing.assert_allclose(out["ou2"], da.rolling(10).corr(db), rtol=1e-6, equal_nan=True)


def test_quantile():  # test_runtime.py:251-268
    f = graph(lambda: Output(WindowedQuantile(Input("a"), 10, 0.49), "ou1"))
    a = np.random.default_rng(3).random((200, 24))
    out = run_oracle(f, {"a": a}, dtype="double", options={"no_fast_stat": "no_warn"})
    exp = pd.DataFrame(a).rolling(10).quantile(0.49, interpolation="linear")
    np.testing.assert_allclose(out["ou1"], exp, rtol=1e-6, equal_nan=True)


def test_large_window_rank_minmax_arg():  # test_runtime.py:271-300 (skip-list path at w=200)
    def build():
        a = Input("a")
        Output(TsRank(a, 200), "ou1")
        Output(WindowedMin(a, 200), "min")
        Output(WindowedMax(a, 200), "max")
        Output(TsArgMin(a, 200), "argmin")
        Output(TsArgMax(a, 200), "argmax")
    f = graph(build)
    a = np.random.default_rng(4).random((2000, 24))
    a[400:410, :] = -1
    out = run_oracle(f, {"a": a}, dtype="double", options={"no_fast_stat": "no_warn"})
    df = pd.DataFrame(a)
    np.testing.assert_allclose(out["ou1"], df.rolling(200).rank(), rtol=1e-6, equal_nan=True)
    np.testing.assert_allclose(out["min"], df.rolling(200).min(), rtol=1e-6, equal_nan=True)
    np.testing.assert_allclose(out["max"], df.rolling(200).max(), rtol=1e-6, equal_nan=True)
    np.testing.assert_allclose(out["argmin"], df.rolling(200).apply(np.argmin) + 1, rtol=1e-6, equal_nan=True)
    np.testing.assert_allclose(out["argmax"], df.rolling(200).apply(np.argmax) + 1, rtol=1e-6, equal_nan=True)


def test_small_window_rank_minmax_arg_with_ties():  # naive loops (w <= 60)
    def build():
        a = Input("a")
        Output(TsRank(a, 7), "rank")
        Output(WindowedMin(a, 7), "min")
        Output(WindowedMax(a, 7), "max")
        Output(TsArgMin(a, 7), "argmin")
        Output(TsArgMax(a, 7), "argmax")
    f = graph(build)
    a = np.floor(np.random.default_rng(5).random((120, 16)) * 4).astype("float32")  # many ties
    a[30, 3] = np.nan
    out = run_oracle(f, {"a": a})
    df = pd.DataFrame(a)
    np.testing.assert_array_equal(out["rank"], df.rolling(7).rank().to_numpy(dtype="float32"))
    np.testing.assert_array_equal(out["min"], df.rolling(7).min().to_numpy(dtype="float32"))
    np.testing.assert_array_equal(out["max"], df.rolling(7).max().to_numpy(dtype="float32"))
    # oldest extreme wins ties, as np.argmin/np.argmax
    exp_min = df.rolling(7).apply(np.argmin, raw=True) + 1
    exp_max = df.rolling(7).apply(np.argmax, raw=True) + 1
    np.testing.assert_array_equal(out["argmin"], exp_min.to_numpy(dtype="float32"))
    np.testing.assert_array_equal(out["argmax"], exp_max.to_numpy(dtype="float32"))


def test_ema_nan_row_and_nested():  # test_runtime.py:304-328
    def build():
        a = Input("a")
        Output(ExpMovingAvg(a, 5), "ou2")
        Output(ExpMovingAvg(ExpMovingAvg(ExpMovingAvg(BackRef(a, 1), 5), 5), 5), "gh_issue_26")
    f = graph(build)
    a = np.random.default_rng(6).random((20, 24)).astype("float32")
    a[5, :] = np.nan
    out = run_oracle(f, {"a": a})
    ref = lambda d: d.ewm(span=5, adjust=False, ignore_na=True).mean()
    df = pd.DataFrame(a)
    np.testing.assert_allclose(out["ou2"], ref(df), rtol=1e-6, equal_nan=True)
    np.testing.assert_allclose(out["gh_issue_26"], ref(ref(ref(df.shift(1)))), rtol=1e-6, equal_nan=True)


def test_ema_init():  # test_runtime.py:332-351
    def build():
        init = Input("__init_1")
        init.attrs["single_value"] = True
        Output(ExpMovingAvg(Input("a"), 5, init), "ou2")
    f = graph(build)
    a = np.random.default_rng(7).random((21, 24)).astype("float32")
    out = run_oracle(f, {"a": a, "__init_1": a[0].copy()}, cur_time=1, length=20)
    exp = pd.DataFrame(a).ewm(span=5, adjust=False, ignore_na=True).mean()
    np.testing.assert_allclose(out["ou2"], exp[1:], rtol=1e-6, equal_nan=True)


def test_argmin_issue19():  # test_runtime.py:355-389
    def build():
        a = Input("a")
        Output(TsArgMin(a, 5), "ou2")
        Output(WindowedMin(a, 5), "tsmin")
        Output(TsRank(a, 5), "tsrank")
    f = graph(build)
    data = [0.6898481863442985, 0.6992020600574415, 0.6992020600574417, 0.6968635916291558,
            0.6968635916291558, 0.6968635916291558]
    a = np.empty((6, 8), "float32")
    for i in range(6):
        a[i, :] = data[i]
    out = run_oracle(f, {"a": a})
    df = pd.DataFrame(a)
    exp = df.rolling(5, min_periods=1).apply(lambda x: x.argmin() + 1, raw=True)
    np.testing.assert_allclose(out["ou2"][4:], exp[4:], rtol=1e-6, equal_nan=True)
    np.testing.assert_allclose(out["tsmin"], df.rolling(5).min(), rtol=1e-6, equal_nan=True)
    np.testing.assert_allclose(out["tsrank"], df.rolling(5).rank(), rtol=1e-6, equal_nan=True)


def test_rank_ties_and_nan_rows():  # test_runtime.py:393-413
    f = graph(lambda: Output(Rank(Input("a")), "ou2"))
    m, p = oracle_module(f)
    run = lambda x: p.run({"a": x}, 24, 20, 0, 20)["ou2"]
    a = np.random.default_rng(8).random((20, 24)).astype("float32")
    exp = lambda x: pd.DataFrame(x).rank(pct=True, axis=1).to_numpy()
    np.testing.assert_allclose(run(a), exp(a), rtol=1e-6, equal_nan=True)
    for i in range(20):
        a[i, :] = i // 2
    np.testing.assert_allclose(run(a), exp(a), rtol=1e-6, equal_nan=True)
    a[10, :] = np.nan
    np.testing.assert_allclose(run(a), exp(a), rtol=1e-6, equal_nan=True)


@pytest.mark.parametrize("stocks", [24, 20])
def test_rank2_bit_exact(stocks):  # test_runtime.py:428-446, rtol=atol=0
    def build():
        a = Input("a")
        v1 = Add(a, a)
        Output(Add(Rank(v1), v1), "out")
    f = graph(build)
    a = np.random.default_rng(9).random((200, stocks))
    out = run_oracle(f, {"a": a}, dtype="double", options={"no_fast_stat": "no_warn"})
    df = pd.DataFrame(a)
    df = df + df
    np.testing.assert_allclose(out["out"], (df.rank(pct=True, axis=1) + df).to_numpy(), rtol=0, atol=0)


def test_rank029_bit_exact():  # test_runtime.py:449-489: FastWindowedSum == pandas rolling sum exactly
    def build():
        a = Input("a")
        inner = WindowedSum(Rank(Rank(-1 * Rank(a))), 5)
        Output(inner, "ou1")
        Output(Rank(inner), "ou2")
    f = graph(build)
    a = np.random.default_rng(10).random((300, 20))
    out = run_oracle(f, {"a": a}, dtype="double", options={"opt_reduce": True, "fast_log": True,
                                                          "no_fast_stat": "no_warn"})
    rank = lambda d: d.rank(axis=1, pct=True)
    df = pd.DataFrame(a)
    inner = rank(rank(-1 * rank(df))).rolling(5).sum()
    np.testing.assert_allclose(out["ou1"], inner, rtol=0, atol=0, equal_nan=True)
    np.testing.assert_allclose(out["ou2"], rank(inner), rtol=0, atol=0, equal_nan=True)


def test_scale():
    f = graph(lambda: Output(Scale(Input("a")), "s"))
    a = np.random.default_rng(11).standard_normal((30, 17))
    a[3, 4] = np.nan
    a[5, :] = 0.0
    out = run_oracle(f, {"a": a}, dtype="double", options={"no_fast_stat": "no_warn"})
    df = pd.DataFrame(a)
    exp = df.div(df.abs().sum(axis=1), axis=0).to_numpy()
    exp[5, :] = np.nan  # 0/0 rows are NaN (Scale.hpp:41)
    np.testing.assert_allclose(out["s"], exp, rtol=1e-12, equal_nan=True)


@pytest.mark.parametrize("dtype", ["float", "double"])
def test_log_special_values(dtype):  # test_runtime.py:501-518 (LogFast)
    npdt = "float32" if dtype == "float" else "float64"
    f = graph(lambda: Output(Log(Input("a")), "outlog"))
    a = np.zeros((20, 24), dtype=npdt)
    for i in range(20):
        a[i, :] = 10.0 ** (i - 10)
    a[0, :] = -10
    a[-1, :] = 0
    a[1, :] = np.nan
    a[2, :] = 1.0
    a[3, :] = np.inf
    out = run_oracle(f, {"a": a}, dtype=dtype)
    with np.errstate(all="ignore"):
        np.testing.assert_allclose(out["outlog"], np.log(a), rtol=1e-5, atol=1e-5, equal_nan=True)
    assert out["outlog"][2, 0] == 0.0 and np.isinf(out["outlog"][-1, 0]) and out["outlog"][-1, 0] < 0


def test_exp_pow():  # test_runtime.py:521-561
    def build():
        a, b = Input("a"), Input("b")
        Output(Pow(a, ConstantOp(0.5)), "sqr")
        Output(Pow(a, ConstantOp(2)), "pow2")
        Output(Pow(a, ConstantOp(5)), "pow5")
        Output(Pow(ConstantOp(1.2), b), "pow12_b")
        Output(Exp(b), "exp")
    f = graph(build)
    base = np.zeros((20, 16), "float32")
    expo = np.zeros((20, 16), "float32")
    for i in range(16):
        base[:, i] = 10.0 ** (i - 8)
        expo[:, i] = 10.0 ** (i / 8 - 1)
    base[:, -1] = 0
    base[:, 1] = np.nan
    expo[:, -1] = 0
    expo[:, 1] = np.nan
    out = run_oracle(f, {"a": base, "b": expo})
    with np.errstate(all="ignore"):
        np.testing.assert_allclose(out["sqr"], np.power(base, 0.5), rtol=1e-5, atol=1e-5, equal_nan=True)
        np.testing.assert_allclose(out["pow2"], np.power(base, 2), rtol=1e-5, atol=1e-5, equal_nan=True)
        np.testing.assert_allclose(out["pow5"], np.power(base, 5), rtol=1e-5, atol=1e-5, equal_nan=True)
        np.testing.assert_allclose(out["pow12_b"], np.power(1.2, expo), rtol=1e-5, atol=1e-5, equal_nan=True)
        np.testing.assert_allclose(out["exp"], np.exp(expo), rtol=1e-5, atol=1e-5, equal_nan=True)


def test_skew_kurt_fast_and_slow():  # test_runtime.py:597-631
    def build():
        a = Input("a")
        Output(WindowedSkew(a, 5), "ou2")
        Output(WindowedKurt(a, 5), "ou3")
        s2, k2 = WindowedSkew(a, 5), WindowedKurt(a, 5)
        s2.attrs["no_fast_stat"] = True
        k2.attrs["no_fast_stat"] = True
        Output(s2, "ou4")
        Output(k2, "ou5")
    f = graph(build)
    a = np.random.default_rng(12).random((20, 24))
    out = run_oracle(f, {"a": a}, dtype="double", options={"no_fast_stat": "no_warn"})
    df = pd.DataFrame(a)
    np.testing.assert_allclose(out["ou2"], df.rolling(5).skew(), rtol=1e-6, equal_nan=True)
    np.testing.assert_allclose(out["ou4"], df.rolling(5).skew(), rtol=1e-6, equal_nan=True)
    np.testing.assert_allclose(out["ou3"], df.rolling(5).kurt(), rtol=1e-6, equal_nan=True)
    np.testing.assert_allclose(out["ou5"], df.rolling(5).kurt(), rtol=1e-6, equal_nan=True)


def test_max_drawdown():  # test_runtime.py:737-771
    f = graph(lambda: Output(WindowedMaxDrawdown(Input("a"), 5), "out"))
    a = np.random.default_rng(13).random((20, 24)).astype("float32")
    out = run_oracle(f, {"a": a})["out"]
    exp = np.full_like(out, np.nan)
    for t in range(4, 20):
        for s in range(24):
            w = a[t - 4:t + 1, s]
            p = int(np.argmax(w))
            exp[t, s] = (w.max() - w[p:].min()) / w.max()
    np.testing.assert_allclose(out[5:], exp[5:], rtol=1e-6, atol=1e-7)


def test_linear_regression():  # test_stream.py:46 (np.polyfit slope) + closed forms for R^2, residual
    def build():
        a = Input("a")
        Output(WindowedLinearRegressionSlope(a, 10), "slope")
        Output(WindowedLinearRegressionRSqaure(a, 10), "rsq")
        Output(WindowedLinearRegressionResi(a, 10), "resi")
    f = graph(build)
    a = np.random.default_rng(14).random((60, 8))
    out = run_oracle(f, {"a": a}, dtype="double", options={"no_fast_stat": "no_warn"})
    x = np.arange(1, 11)
    for t in range(9, 60):
        for s in range(8):
            y = a[t - 9:t + 1, s]
            k, c = np.polyfit(x, y, 1)
            r2 = np.corrcoef(x, y)[0, 1] ** 2
            np.testing.assert_allclose(out["slope"][t, s], k, rtol=1e-8)
            np.testing.assert_allclose(out["rsq"][t, s], r2, rtol=1e-8)
            np.testing.assert_allclose(out["resi"][t, s], y[-1] - (k * 10 + c), rtol=1e-6, atol=1e-9)


def test_decay_linear_closed_form():
    f = graph(lambda: Output(DecayLinear(Input("a"), 6), "d"))
    a = np.random.default_rng(15).random((40, 8))
    out = run_oracle(f, {"a": a}, dtype="double", options={"no_fast_stat": "no_warn"})
    w = np.arange(1, 7) / (6 * 7 / 2)
    for t in range(5, 40):
        np.testing.assert_allclose(out["d"][t], (a[t - 5:t + 1] * w[:, None]).sum(0), rtol=1e-12)
    assert np.isnan(out["d"][:5]).all()


def test_warmup_masks_follow_unreliable_counts():
    def build():
        a = Input("a")
        Output(WindowedSum(WindowedAvg(a, 5), 7), "s")
        Output(TsRank(BackRef(a, 3), 4), "r")
    f = graph(build)
    m, p = oracle_module(f)
    a = np.random.default_rng(16).random((50, 8)).astype("float32")
    out = p.run({"a": a}, 8, 50, 0, 50)
    # first valid row = sum of (window - 1) along the chain; BackRef(w) needs w rows
    first_valid = {"s": (5 - 1) + (7 - 1), "r": 3 + (4 - 1)}
    for b in m.buffers:
        if b.kind == "OUTPUT":
            o = out[b.name]
            fv = first_valid[b.name]
            assert np.isnan(o[:fv]).all() and not np.isnan(o[fv:]).any()
            assert fv <= b.unreliable
    assert {b.name: b.unreliable for b in m.buffers if b.kind == "OUTPUT"} == {"s": 12, "r": 7}


def test_no_history_before_cur_time():  # Ops.hpp:87-92: rows before cur_time are never read
    f = graph(lambda: Output(WindowedSum(Input("a"), 5), "s"))
    a = np.random.default_rng(17).random((30, 8))
    full = run_oracle(f, {"a": a}, dtype="double", options={"no_fast_stat": "no_warn"})
    part = run_oracle(f, {"a": a}, cur_time=10, length=20, dtype="double", options={"no_fast_stat": "no_warn"})
    fresh = run_oracle(f, {"a": a[10:].copy()}, dtype="double", options={"no_fast_stat": "no_warn"})
    np.testing.assert_array_equal(part["s"], fresh["s"])
    assert np.isnan(part["s"][:4]).all() and not np.isnan(full["s"][4:14]).any()


def test_stream_matches_pandas():  # test_stream.py:12-47
    from oracle.kun_oracle import OracleStream
    def build():
        a = Input("a")
        Output(WindowedQuantile(a, 10, 0.49), "quantile")
        Output(ExpMovingAvg(a, 10), "ema")
        Output(WindowedLinearRegressionSlope(a, 10), "slope")
    f = graph(build)
    m, p = oracle_module(f, dtype="double", input_layout="STREAM", output_layout="STREAM")
    st = OracleStream(p, 24)
    a = np.random.default_rng(18).random((100, 24))
    q, e, sl = (np.empty((100, 24)) for _ in range(3))
    for i in range(100):
        st.push("a", a[i])
        st.run()
        q[i], e[i], sl[i] = st.get("quantile"), st.get("ema"), st.get("slope")
    df = pd.DataFrame(a)
    np.testing.assert_allclose(q, df.rolling(10).quantile(0.49, interpolation="linear"), atol=1e-6, rtol=1e-4, equal_nan=True)
    np.testing.assert_allclose(e, df.ewm(span=10, adjust=False, ignore_na=True).mean(), atol=1e-6, rtol=1e-4, equal_nan=True)
    exp = df.rolling(10).apply(lambda x: np.polyfit(np.arange(len(x)), x, 1)[0]).to_numpy()
    np.testing.assert_allclose(sl[10:], exp[10:], atol=1e-6, rtol=1e-4, equal_nan=True)


def test_stream_equals_batch():
    """A stream of T ticks reproduces the batch result of the same graph."""
    from oracle.kun_oracle import OracleStream
    def build():
        a = Input("a")
        Output(WindowedStddev(a, 6), "std")
        Output(TsRank(a, 4), "tsr")
        Output(Rank(DecayLinear(a, 3)), "rk")
    f1, f2 = graph(build), graph(build)
    a = np.random.default_rng(19).random((40, 16))
    batch = run_oracle(f1, {"a": a}, dtype="double", options={"no_fast_stat": True})
    m, p = oracle_module(f2, dtype="double", input_layout="STREAM", output_layout="STREAM")
    st = OracleStream(p, 16)
    for i in range(40):
        st.push("a", a[i])
        st.run()
        for k in ("std", "tsr", "rk"):
            np.testing.assert_array_equal(st.get(k), batch[k][i])


# ---- corrWith / aggregrate (test_runtime.py:17-122) ------------------------

def test_oracle_corrwith_matches_pandas():
    """CorrWith.hpp against pandas corrwith(axis=1), plain and on ranks, at
    the reference test's tolerance; TS and STs layouts agree exactly."""
    import pandas as pd
    from oracle import kun_oracle as ko
    rng = np.random.default_rng(41)
    a = rng.random((24, 20)).astype("float32")
    r = rng.random((24, 20)).astype("float32")
    np.testing.assert_allclose(ko.corr_with(a, r), pd.DataFrame(a).corrwith(pd.DataFrame(r), axis=1),
                               atol=1e-6, rtol=1e-4)
    a2 = pd.DataFrame(a).rank(axis=1, pct=True).astype("float32")
    np.testing.assert_allclose(ko.corr_with(a, r, rank_inputs=True), a2.corrwith(pd.DataFrame(r), axis=1),
                               atol=1e-6, rtol=1e-4)
    sts = lambda x: np.ascontiguousarray(x.reshape(24, 20 // 4, 4).transpose(1, 0, 2))
    assert np.array_equal(ko.corr_with(sts(a), sts(r), layout="STs"), ko.corr_with(a, r), equal_nan=True)
    # a NaN cell makes its row NaN (no NaN skipping in the sums)
    a[3, 7] = np.nan
    out = ko.corr_with(a, r)
    assert np.isnan(out[3]) and not np.isnan(out[4])


@pytest.mark.parametrize("dtype", ["float32", "float64"])
def test_oracle_aggregate_matches_pandas(dtype):
    """Aggregration.cpp against pandas groupby(...).sum/min/max/first/last/
    count/mean over runs of equal labels (the reference test's data shape)."""
    import pandas as pd
    from oracle import kun_oracle as ko
    rng = np.random.default_rng(42)
    a = rng.random((240, 16)).astype(dtype)
    dates = pd.date_range(start="2026-01-01", end="2026-01-07", freq="min")
    stamps = pd.DatetimeIndex(sorted(rng.choice(dates, size=240)))
    df = pd.DataFrame(a, index=stamps)
    g = df.groupby(df.index.date)
    labels = stamps.day.to_numpy().astype(dtype)
    got = ko.aggregate(a, labels)
    for k in ko.AGG_KINDS:
        np.testing.assert_allclose(got[k], getattr(g, k)(), atol=1e-6, rtol=1e-4, err_msg=k)
    # start offset: the groups of rows 30.. only
    sub = ko.aggregate(a, labels, start=30)
    g2 = pd.DataFrame(a[30:], index=stamps[30:]).groupby(stamps[30:].date)
    np.testing.assert_allclose(sub["sum"], g2.sum(), atol=1e-6, rtol=1e-4)


def test_diff_with_weighted_sum():  # test_runtime.py:78-99 (atol=1e-6, rtol=1e-4)
    from kunquant_amd.Driver import compileit
    from kunquant_amd.codegen.program import program_text
    from kunquant_amd.workloads.graphs import generic_cs_graph
    from oracle.kun_oracle import OracleProgram
    rng = np.random.default_rng(9)
    ins = {k: rng.random((24, 16)).astype(np.float32) for k in ("a", "b", "c")}
    ins["a"][3, 4] = np.nan
    prog = OracleProgram(program_text(compileit(generic_cs_graph(), "g", input_layout="TS", output_layout="TS")))
    got = prog.run(ins, 16, 24, 0, 24)
    v1 = (ins["a"] - ins["b"]).astype(np.float64)
    exp = v1 - np.sum(v1 * ins["c"], axis=1).reshape(-1, 1)
    np.testing.assert_allclose(got["out"], exp, atol=1e-6, rtol=1e-4, equal_nan=True)
    assert np.isnan(got["out"][3]).all() and not np.isnan(np.delete(got["out"], 3, axis=0)).any()
