"""Factor graphs of the benchmark workloads and the test catalog.

Every builder takes a namespace `K` of graph constructors (Builder, Input,
Output, op classes). By default it is this package's front-end; the golden
vector script passes the reference's classes so that both compilers see the
same graph.

simple_graph: x = (close - open) / ((high - low) + 0.001), WindowedStddev(x, 20),
WindowedSum(volume, 10) and the cross-sectional Rank of x.

random_graph(n, seed): n seeded random expression trees of depth 3-8 over
open/high/low/close/volume/amount with elementwise arithmetic,
compare + Select, Delay/Delta, WindowedSum/Avg/Stddev/Corr/Cov,
TsArgMax/TsRank and DecayLinear (windows from {5, 10, 20, 30, 60}); about
30% of the internal nodes are cross-sectional Rank or Scale.

precision_graph(n, seed): the same trees plus WindowedQuantile, windowed
linear-regression slope / R-squared and 60-row variance / correlation /
covariance, 10% cross-sectional nodes."""
import random
import types

FIELDS = ("open", "high", "low", "close", "volume", "amount")
WINDOWS = (5, 10, 20, 30, 60)

_NAMES = ["Builder", "Input", "Output", "Rank", "Scale", "ConstantOp", "Function", "Abs", "Add", "BackRef",
          "DecayLinear", "Div", "GreaterThan", "LessThan", "Mul", "Select", "Sub", "TsArgMax", "TsArgMin",
          "TsRank", "WindowedAvg", "WindowedCorrelation", "WindowedCovariance", "WindowedStddev", "WindowedSum",
          "WindowedMin", "WindowedMax", "ExpMovingAvg", "WindowedQuantile", "WindowedLinearRegressionSlope",
          "WindowedLinearRegressionRSqaure", "WindowedLinearRegressionResi", "WindowedSkew", "WindowedKurt",
          "WindowedMaxDrawdown", "Log", "Exp", "Sqrt", "Sign", "SetInfOrNanToValue", "Max", "Min", "Pow", "Clip",
          "And", "Or", "Not", "LessEqual", "Equals", "WindowedVar", "DiffWithWeightedSum"]


def namespace(op_mod, ops_mod, stage_mod):
    """Collect constructors from (Op, ops, Stage) modules of a front-end."""
    ns = {}
    for n in _NAMES:
        for m in (op_mod, ops_mod, stage_mod):
            if hasattr(m, n):
                ns[n] = getattr(m, n)
                break
    return types.SimpleNamespace(**ns)


def native():
    from .. import Op, Stage, ops
    return namespace(Op, ops, Stage)


def simple_graph(K=None):
    K = K or native()
    b = K.Builder()
    with b:
        close, opn, high, low, vol = (K.Input(n) for n in ("close", "open", "high", "low", "volume"))
        x = (close - opn) / ((high - low) + 0.001)
        K.Output(x, "x")
        K.Output(K.WindowedStddev(x, 20), "std20")
        K.Output(K.WindowedSum(vol, 10), "vsum10")
        K.Output(K.Rank(x), "rank_x")
    return K.Function(b.ops)


class _TreeGen:
    def __init__(self, K, seed: int, cs_frac: float = 0.3):
        self.K = K
        self.rng = random.Random(seed)
        self.cs_frac = cs_frac
        self.inputs = {}
        self.counts = {"nodes": 0, "cs": 0, "windowed": 0}

    def leaf(self):
        name = self.rng.choice(FIELDS)
        if name not in self.inputs:
            self.inputs[name] = self.K.Input(name)
        return self.inputs[name]

    def window(self) -> int:
        return self.rng.choice(WINDOWS)

    def tree(self, depth: int):
        K, r = self.K, self.rng
        if depth <= 0:
            return self.leaf()
        self.counts["nodes"] += 1
        if r.random() < self.cs_frac:
            self.counts["cs"] += 1
            child = self.tree(depth - 1)
            return K.Rank(child) if r.random() < 0.6 else K.Scale(child)
        kind = r.random()
        if kind < 0.22:
            a = self.tree(depth - 1)
            b = self.tree(r.randint(0, depth - 1))
            if r.random() < 0.5:
                a, b = b, a
            op = r.choice((K.Add, K.Sub, K.Mul, K.Div))
            return op(a, b)
        if kind < 0.30:
            a = self.tree(depth - 1)
            c = r.choice((0.5, 2.0, 1.5, -1.0, 10.0))
            choice = r.randrange(3)
            return a * c if choice == 0 else (a + c if choice == 1 else K.Abs(a))
        if kind < 0.36:
            a = self.tree(depth - 1)
            b = self.tree(r.randint(0, depth - 1))
            cmp = K.GreaterThan(a, b) if r.random() < 0.5 else K.LessThan(a, b)
            return K.Select(cmp, a, b)
        if kind < 0.44:
            self.counts["windowed"] += 1
            a = self.tree(depth - 1)
            w = self.window()
            return K.BackRef(a, w) if r.random() < 0.5 else K.Sub(a, K.BackRef(a, w))
        if kind < 0.88:
            self.counts["windowed"] += 1
            a = self.tree(depth - 1)
            op = r.choice((K.WindowedSum, K.WindowedAvg, K.WindowedStddev, K.TsArgMax, K.TsRank, K.DecayLinear))
            return op(a, self.window())
        self.counts["windowed"] += 1
        a = self.tree(depth - 1)
        b = self.tree(r.randint(0, depth - 1))
        op = K.WindowedCorrelation if r.random() < 0.5 else K.WindowedCovariance
        return op(a, self.window(), b)


class _PrecisionTreeGen(_TreeGen):
    """Trees of the precision workload: the random-graph op mix plus
    WindowedQuantile (q = 0.2, 0.8), windowed linear-regression slope and
    R-squared, and long-window (60) variance and correlation, with fewer
    cross-sectional nodes."""

    def tree(self, depth: int):
        K, r = self.K, self.rng
        if depth <= 0:
            return self.leaf()
        if r.random() < 0.3:
            self.counts["nodes"] += 1
            self.counts["windowed"] += 1
            a = self.tree(depth - 1)
            kind = r.randrange(6)
            if kind == 0:
                return K.WindowedQuantile(a, self.window(), r.choice((0.2, 0.8)))
            if kind == 1:
                return K.WindowedLinearRegressionSlope(a, self.window())
            if kind == 2:
                return K.WindowedLinearRegressionRSqaure(a, self.window())
            if kind == 3:
                return K.WindowedStddev(a, 60)
            if kind == 4:
                return K.WindowedCorrelation(a, 60, self.tree(r.randint(0, depth - 1)))
            return K.WindowedCovariance(a, 60, self.tree(r.randint(0, depth - 1)))
        return super().tree(depth)


def precision_graph(num_factors: int, seed: int, prefix: str = "p", K=None):
    """Seeded factor trees of depth 3-8 for the precision run (see
    _PrecisionTreeGen); inputs are the OHLCV fields."""
    K = K or native()
    gen = _PrecisionTreeGen(K, seed, cs_frac=0.1)
    b = K.Builder()
    with b:
        for i in range(num_factors):
            K.Output(gen.tree(gen.rng.randint(3, 8)), f"{prefix}{i:03d}")
    f = K.Function(b.ops)
    f.workload_stats = dict(gen.counts)
    return f


def random_graph(num_factors: int, seed: int, prefix: str = "f", K=None):
    K = K or native()
    gen = _TreeGen(K, seed)
    b = K.Builder()
    with b:
        for i in range(num_factors):
            depth = gen.rng.randint(3, 8)
            K.Output(gen.tree(depth), f"{prefix}{i:03d}")
    f = K.Function(b.ops)
    f.workload_stats = dict(gen.counts)
    return f


def ops_graph(K=None):
    """One output per primitive family, over inputs a and b."""
    K = K or native()
    b = K.Builder()
    with b:
        a, c = K.Input("a"), K.Input("b")
        K.Output(K.WindowedSum(a, 10), "sum10")
        K.Output(K.WindowedAvg(a, 7), "avg7")
        K.Output(K.WindowedStddev(a, 20), "std20")
        K.Output(K.WindowedCovariance(a, 10, c), "cov10")
        K.Output(K.WindowedCorrelation(a, 10, c), "corr10")
        K.Output(K.WindowedMin(a, 5), "min5")
        K.Output(K.WindowedMax(a, 5), "max5")
        K.Output(K.TsArgMax(a, 5), "argmax5")
        K.Output(K.TsArgMin(a, 6), "argmin6")
        K.Output(K.TsRank(a, 7), "tsrank7")
        K.Output(K.DecayLinear(a, 9), "decay9")
        K.Output(K.ExpMovingAvg(a, 5), "ema5")
        K.Output(K.WindowedQuantile(a, 10, 0.49), "q49")
        K.Output(K.WindowedQuantile(a, 12, 0.8), "q80")
        K.Output(K.WindowedLinearRegressionSlope(a, 10), "slope10")
        K.Output(K.WindowedLinearRegressionRSqaure(a, 10), "rsq10")
        K.Output(K.WindowedLinearRegressionResi(a, 10), "resi10")
        K.Output(K.WindowedSkew(a, 5), "skew5")
        K.Output(K.WindowedKurt(a, 6), "kurt6")
        K.Output(K.WindowedMaxDrawdown(a, 5), "mdd5")
        K.Output(K.BackRef(a, 3), "delay3")
        K.Output(a - K.BackRef(a, 2), "delta2")
        K.Output(K.Rank(a), "rank")
        K.Output(K.Scale(c), "scale")
        K.Output(K.Log(a), "log")
        K.Output(K.Exp(c), "exp")
        K.Output(K.Sqrt(a), "sqrt")
        K.Output(K.Sign(a - c), "sign")
        K.Output(K.SetInfOrNanToValue(a / c, 7.0), "finite")
        K.Output(K.Select(K.GreaterThan(a, c), a, c * 2), "select")
        K.Output(K.Max(a, c) - K.Min(a, c), "maxmin")
        K.Output(K.Pow(a, K.ConstantOp(3)), "pow3")
        K.Output(K.Pow(a, K.ConstantOp(0.5)), "pow05")
        K.Output(K.Clip(a - c, 0.3), "clip")
        K.Output(K.Select(K.And(K.Not(K.LessEqual(a, c)), K.Or(K.Equals(a, a), K.Equals(c, c))), a, c), "logic")
    return K.Function(b.ops)


def stream_graph(K=None):
    """Stream-mode checks: quantile, EMA, regression slope, window stats and
    a cross-sectional rank."""
    K = K or native()
    b = K.Builder()
    with b:
        a = K.Input("a")
        K.Output(K.WindowedQuantile(a, 10, 0.49), "quantile")
        K.Output(K.ExpMovingAvg(a, 10), "ema")
        K.Output(K.WindowedLinearRegressionSlope(a, 10), "slope")
        K.Output(K.WindowedStddev(a, 10), "std")
        K.Output(K.TsRank(a, 5), "tsrank")
        K.Output(K.Rank(K.WindowedSum(a, 5)), "rank")
    return K.Function(b.ops)


def capi_batch_graph(K=None):
    """output = input + input * 2 over one input (the module the reference's
    C API test runs in STs layout)."""
    K = K or native()
    b = K.Builder()
    with b:
        inp = K.Input("input")
        K.Output(inp + inp * 2, "output")
    return K.Function(b.ops)


def capi_stream_graph(K=None):
    """alpha101 = (close - open) / ((high - low) + 0.001) beside windowed
    factors over the six fields, so that a stream's serialised state holds
    rings and op state (the stream module of the reference's C API test)."""
    K = K or native()
    b = K.Builder()
    with b:
        close, opn, high, low, vol, amount = (K.Input(n) for n in FIELDS[3:4] + FIELDS[0:1] + FIELDS[1:3]
                                              + FIELDS[4:6])
        K.Output((close - opn) / ((high - low) + 0.001), "alpha101")
        K.Output(K.WindowedCorrelation(close, 10, vol), "corr10")
        K.Output(K.ExpMovingAvg(amount, 5), "ema5")
        K.Output(K.Rank(K.WindowedSum(high - low, 5)), "rank_range5")
    return K.Function(b.ops)


def aligned_graph(K=None):
    """Rank(a) * 2; compiled with allow_unaligned=False."""
    K = K or native()
    b = K.Builder()
    with b:
        K.Output(K.Rank(K.Input("a")) * 2, "ou2")
    return K.Function(b.ops)


ALPHA158_CONFIG = {
    "kbar": {},
    "price": {"windows": [0], "feature": ["OPEN", "HIGH", "LOW", "VWAP"]},
    "rolling": {"windows": [5, 10, 20, 30, 60]},
}


def alpha158_graph(K=None, AllData=None):
    """The Alpha158 set as the reference's test_alpha158.py:17-38 builds it:
    kbar features, OPEN/HIGH/LOW/VWAP of the day over close, and every
    rolling feature over windows 5, 10, 20, 30, 60 (158 outputs). AllData:
    the factor library's class (this package's by default)."""
    K = K or native()
    if AllData is None:
        from ..predefined.Alpha158 import AllData
    b = K.Builder()
    with b:
        pack = AllData(low=K.Input("low"), high=K.Input("high"), close=K.Input("close"), open=K.Input("open"),
                       amount=K.Input("amount"), volume=K.Input("volume"))
        cfg = {k: dict(v) for k, v in ALPHA158_CONFIG.items()}
        cfg["price"]["feature"] = [(n, getattr(pack, n.lower())) for n in ALPHA158_CONFIG["price"]["feature"]]
        fields, names = pack.build(cfg)
        for v, k in zip(fields, names):
            K.Output(v, k)
    return K.Function(b.ops)


def large_window_graph(K=None):
    """TsRank, WindowedMin/Max, TsArgMin/TsArgMax over 200 rows: windows
    long enough for the reference's skip-list form in fp64
    (CompOp.py:14-17)."""
    K = K or native()
    b = K.Builder()
    with b:
        a = K.Input("a")
        K.Output(K.TsRank(a, 200), "rank200")
        K.Output(K.WindowedMin(a, 200), "min200")
        K.Output(K.WindowedMax(a, 200), "max200")
        K.Output(K.TsArgMin(a, 200), "argmin200")
        K.Output(K.TsArgMax(a, 200), "argmax200")
    return K.Function(b.ops)


def alpha101_graph(K=None, library=None):
    """Every alpha of an Alpha101 library (`library`: a module with
    AllData and all_alpha; default this package's) over the six fields."""
    K = K or native()
    if library is None:
        from ..predefined import Alpha101 as library
    b = K.Builder()
    with b:
        d = library.AllData(low=K.Input("low"), high=K.Input("high"), close=K.Input("close"), open=K.Input("open"),
                            amount=K.Input("amount"), volume=K.Input("volume"))
        for f in library.all_alpha:
            K.Output(f(d), f.__name__)
    return K.Function(b.ops)


def tsrank_chain_graph(K=None):
    """A factor of the 100-factor graph (f022) with its intermediates as
    outputs: TsArgMax over TsRank over TsRank over a rolling sum of the
    cross-sectional rank of volume. The second TsRank's current value is the
    first TsRank, the case where the reference's rank continues the other
    rank's counts."""
    K = K or native()
    b = K.Builder()
    with b:
        r1 = K.Rank(K.Input("volume"))
        s = K.Scale(r1)
        r2 = K.Rank(s)
        w = K.WindowedSum(r2, 10)
        t60 = K.TsRank(w, 60)
        t30 = K.TsRank(t60, 30)
        K.Output(r2, "r2")
        K.Output(w, "wsum")
        K.Output(t60, "t60")
        K.Output(t30, "t30")
        K.Output(K.Abs(K.TsArgMax(t30, 20)), "f022")
    return K.Function(b.ops)


def generic_cs_graph(K=None):
    """A generic cross-sectional op (body compiled as device code) between
    time-series stages: DiffWithWeightedSum(a - b, c) (the reference's
    test_runtime.py:78-99 graph), its 5-row average and its rank."""
    K = K or native()
    b = K.Builder()
    with b:
        a, bb, c = K.Input("a"), K.Input("b"), K.Input("c")
        d = K.DiffWithWeightedSum(a - bb, c)
        K.Output(d, "out")
        K.Output(K.WindowedAvg(d, 5), "avg5")
        K.Output(K.Rank(d), "rank")
    return K.Function(b.ops)


def cross_section_graph(K=None):
    """Rank and Scale straight off the inputs (cross-sectional kernel checks)."""
    K = K or native()
    b = K.Builder()
    with b:
        a = K.Input("a")
        c = K.Input("b")
        K.Output(K.Rank(a), "rank_a")
        K.Output(K.Scale(a), "scale_a")
        K.Output(K.Rank(c), "rank_b")
        K.Output(K.Scale(c), "scale_b")
    return K.Function(b.ops)
