"""Alpha158 factor set (the qlib Alpha158 handler's feature list) as a graph
of front-end ops, with the interface of KunQuant/predefined/Alpha158.py:6-242
(`AllData(...)`, `build(config) -> (fields, names)`, the same config keys,
feature names and order).

Formulas follow qlib's Alpha158 definitions with the reference's choices,
which fix the numbers: vwap = amount / (volume + 1e-6); RANK = TsRank / d;
IMAX / IMIN via TsArgMax / TsArgMin divided by d; RSV clipped to +-10; CORR
and CORD with inf/NaN set to 0 and CORD's volume ratio taken over
(Ref(volume, 1) + 1e-12); up/down flags as Select(cond, 1, 0); the MIN
features are switched by the key "LOW" (Alpha158.py:114-118). Every
expression keeps the reference's operand order (Max/Min pick their second
operand on NaN, so order is part of the result)."""
from ..Op import ConstantOp, OpBase
from ..ops import (Abs, BackRef, Clip, Log, Max, Min, Select, SetInfOrNanToValue, TsArgMax, TsArgMin, TsRank,
                   WindowedAvg, WindowedCorrelation, WindowedLinearRegressionResi,
                   WindowedLinearRegressionRSqaure, WindowedLinearRegressionSlope, WindowedMax, WindowedMin,
                   WindowedQuantile, WindowedStddev, WindowedSum)

EPS = 1e-12


def _flag(cond: OpBase) -> OpBase:
    return Select(cond, ConstantOp(1), ConstantOp(0))


class AllData:
    def __init__(self, open: OpBase, close: OpBase = None, high: OpBase = None, low: OpBase = None,
                 volume: OpBase = None, amount: OpBase = None, vwap: OpBase = None) -> None:
        self.open, self.close, self.high, self.low = open, close, high, low
        self.volume, self.amount = volume, amount
        self.vwap = vwap if vwap is not None else self.amount / (self.volume + 0.000001)

    # ---- building blocks ------------------------------------------------
    def _range(self) -> OpBase:
        return self.high - self.low + EPS

    def _moves(self, x: OpBase):
        """(up move, down move, |move|) of x against the previous row."""
        up = Max(x - BackRef(x, 1), ConstantOp(0))
        down = Max(BackRef(x, 1) - x, ConstantOp(0))
        return up, down, Abs(x - BackRef(x, 1))

    def _kbar(self):
        o, c, h, lo = self.open, self.close, self.high, self.low
        body_top, body_bottom = Max(o, c), Min(o, c)
        shift = 2 * c - h - lo
        return [("KMID", (c - o) / o), ("KLEN", (h - lo) / o), ("KMID2", (c - o) / self._range()),
                ("KUP", (h - body_top) / o), ("KUP2", (h - body_top) / self._range()),
                ("KLOW", (body_bottom - lo) / o), ("KLOW2", (body_bottom - lo) / self._range()),
                ("KSFT", shift / o), ("KSFT2", shift / self._range())]

    def _rolling_table(self):
        """[(config key, name prefix, factor of window d)] in qlib's order."""
        c, v, h, lo = self.close, self.volume, self.high, self.low

        def move_ratio(x, part):
            def f(d):
                up, down, mag = self._moves(x)
                num = {"P": lambda: WindowedSum(up, d), "N": lambda: WindowedSum(down, d),
                       "D": lambda: WindowedSum(up, d) - WindowedSum(self._moves(x)[1], d)}[part]()
                return num / (WindowedSum(mag, d) + EPS)
            return f

        def wvma(d):
            w = Abs(c / BackRef(c, 1) - 1) * v
            return WindowedStddev(w, d) / (WindowedAvg(Abs(c / BackRef(c, 1) - 1) * v, d) + EPS)

        up_day = lambda: _flag(c > BackRef(c, 1))
        down_day = lambda: _flag(c < BackRef(c, 1))
        return [
            ("ROC", "ROC", lambda d: BackRef(c, d) / c),
            ("MA", "MA", lambda d: WindowedAvg(c, d) / c),
            ("STD", "STD", lambda d: WindowedStddev(c, d) / c),
            ("BETA", "BETA", lambda d: WindowedLinearRegressionSlope(c, d) / c),
            ("RSQR", "RSQR", lambda d: WindowedLinearRegressionRSqaure(c, d)),
            ("RESI", "RESI", lambda d: WindowedLinearRegressionResi(c, d) / c),
            ("MAX", "MAX", lambda d: WindowedMax(h, d) / c),
            ("LOW", "MIN", lambda d: WindowedMin(lo, d) / c),
            ("QTLU", "QTLU", lambda d: WindowedQuantile(c, d, 0.8) / c),
            ("QTLD", "QTLD", lambda d: WindowedQuantile(c, d, 0.2) / c),
            ("RANK", "RANK", lambda d: TsRank(c, d) / d),
            ("RSV", "RSV", lambda d: Clip((c - WindowedMin(lo, d)) / (WindowedMax(h, d) - WindowedMin(lo, d) + EPS),
                                          10)),
            ("IMAX", "IMAX", lambda d: TsArgMax(h, d) / d),
            ("IMIN", "IMIN", lambda d: TsArgMin(lo, d) / d),
            ("IMXD", "IMXD", lambda d: (TsArgMax(h, d) - TsArgMin(lo, d)) / d),
            ("CORR", "CORR", lambda d: SetInfOrNanToValue(WindowedCorrelation(c, d, Log(v + 1)))),
            ("CORD", "CORD", lambda d: SetInfOrNanToValue(
                WindowedCorrelation(c / BackRef(c, 1), d, Log(v / (BackRef(v, 1) + EPS) + 1)))),
            ("CNTP", "CNTP", lambda d: WindowedAvg(up_day(), d)),
            ("CNTN", "CNTN", lambda d: WindowedAvg(down_day(), d)),
            ("CNTD", "CNTD", lambda d: WindowedAvg(up_day(), d) - WindowedAvg(down_day(), d)),
            ("SUMP", "SUMP", move_ratio(c, "P")),
            ("SUMN", "SUMN", move_ratio(c, "N")),
            ("SUMD", "SUMD", move_ratio(c, "D")),
            ("VMA", "VMA", lambda d: WindowedAvg(v, d) / (v + EPS)),
            ("VSTD", "VSTD", lambda d: WindowedStddev(v, d) / (v + EPS)),
            ("WVMA", "WVMA", wvma),
            ("VSUMP", "VSUMP", move_ratio(v, "P")),
            ("VSUMN", "VSUMN", move_ratio(v, "N")),
            ("VSUMD", "VSUMD", move_ratio(v, "D")),
        ]

    def build(self, config: dict):
        """Factors and their names for `config`:
            {'kbar': {}, 'price': {'windows': [0, 1, ...], 'feature': [(name, op), ...]},
             'volume': {'windows': [...]},
             'rolling': {'windows': [5, 10, 20, 30, 60], 'include': [...] or None, 'exclude': [...]}}
        (a section present enables it; defaults as in the reference)."""
        out = []
        if "kbar" in config:
            out += self._kbar()
        if "price" in config:
            windows = config["price"].get("windows", range(5))
            feature = config["price"].get("feature", [("OPEN", self.open), ("HIGH", self.high), ("LOW", self.low)])
            for fname, x in feature:
                out += [(fname.upper() + str(d), (BackRef(x, d) if d else x) / self.close) for d in windows]
        if "volume" in config:
            windows = config["volume"].get("windows", range(5))
            out += [("VOLUME" + str(d), (BackRef(self.volume, d) if d else self.volume) / (self.volume + EPS))
                    for d in windows]
        if "rolling" in config:
            windows = config["rolling"].get("windows", [5, 10, 20, 30, 60])
            include = config["rolling"].get("include", None)
            exclude = config["rolling"].get("exclude", [])
            for key, prefix, make in self._rolling_table():
                if key in exclude or (include is not None and key not in include):
                    continue
                out += [("%s%d" % (prefix, d), make(d)) for d in windows]
        return [f for _, f in out], [n for n, _ in out]
