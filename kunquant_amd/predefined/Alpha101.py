"""Alpha101 factor set ("101 Formulaic Alphas", Kakushadze 2015) as a graph
of front-end ops, with the interface of KunQuant/predefined/Alpha101.py
(`AllData(...)`, the per-alpha builder functions `alpha001` ... `alpha101`
and their list `all_alpha`, same names and order, lines 541-548).

Each alpha is written as one formula in a small expression language (the
paper's notation) and built by the parser below. The formulas carry the
reference's formulation choices, which fix its numbers: integer windows
(the paper's fractional windows rounded as the reference rounds them),
inf/NaN replaced by a value around correlations and ratios (`fill`), zero
divisors replaced by 0.0001, booleans as 1/0 (`b10`), two-pass window
statistics where the reference asks for them (`correlation_twopass`,
`sma_twopass`), a few deviations from the paper (alpha001 without the 0.5
shift, alpha038 ranking open, alpha039 with a 250-day mean of returns,
alpha052 as a delta of ts_min, ...) and the operand order and grouping of
every expression (Min/Max pick their second operand on NaN; rounding
follows the grouping).

Language: numbers; the fields open, close, high, low, volume, amount,
vwap, returns; + - * / with the usual precedence; comparisons < <= > >=
== (binding tighter than & and |, as in C); & | ~ on conditions; the
ternary c ? a : b (Select); and the functions of _FUNCS. Arithmetic on two
numbers is folded (as Python folds the reference's `1 - 0.178404`)."""
import re
from typing import Callable, Dict, List, Tuple

from ..Op import ConstantOp, OpBase, Rank, Scale
from ..ops import (Abs, BackRef, DecayLinear, Equals, Log, Max, Min, Pow, Select, SetInfOrNanToValue, Sign, TsArgMax,
                   TsArgMin, TsRank, WindowedAvg, WindowedCorrelation, WindowedCovariance, WindowedMax, WindowedMin,
                   WindowedProduct, WindowedStddev, WindowedSum)


class AllData:
    def __init__(self, open: OpBase, close: OpBase = None, high: OpBase = None, low: OpBase = None,
                 volume: OpBase = None, amount: OpBase = None, vwap: OpBase = None) -> None:
        self.open, self.close, self.high, self.low = open, close, high, low
        self.volume, self.amount = volume, amount
        self.vwap = vwap if vwap is not None else amount / (volume + 0.0000001)
        self.returns = close / BackRef(close, 1) - 1.0


# ---- the formulas ----------------------------------------------------------
# (name, formula); `adv(n)` is the n-day mean volume.
_FORMULAS: List[Tuple[str, str]] = [
    ("alpha001", "rank(ts_argmax(sq(returns < 0 ? stddev(returns, 20) : close), 5))"),
    ("alpha002", "fill(correlation(rank(delta(log(volume), 2)), rank((close - open) / open), 6) * -1)"),
    ("alpha003", "fill(correlation(rank(open), rank(volume), 10) * -1)"),
    ("alpha004", "ts_rank(rank(low), 9) * -1"),
    ("alpha005", "rank(open - ts_sum(vwap, 10) / 10) * (abs(rank(close - vwap)) * -1)"),
    ("alpha006", "fill(correlation(open, volume, 10) * -1)"),
    ("alpha007", "(adv(20) >= volume) ? -1 : ((ts_rank(abs(delta(close, 7)), 60) * sign(delta(close, 7))) * -1)"),
    ("alpha008", "rank(ts_sum(open, 5) * ts_sum(returns, 5) - delay(ts_sum(open, 5) * ts_sum(returns, 5), 10)) * -1"),
    ("alpha009", "((ts_min(delta(close, 1), 5) > 0) | (ts_max(delta(close, 1), 5) < 0)) ? delta(close, 1)"
                 " : (delta(close, 1) * -1)"),
    ("alpha010", "((ts_min(delta(close, 1), 4) > 0) | (ts_max(delta(close, 1), 4) < 0)) ? delta(close, 1)"
                 " : (delta(close, 1) * -1)"),
    ("alpha011", "(rank(ts_max(vwap - close, 3)) + rank(ts_min(vwap - close, 3))) * rank(delta(volume, 3))"),
    ("alpha012", "sign(delta(volume, 1)) * (-1 * delta(close, 1))"),
    ("alpha013", "-1 * rank(covariance(rank(close), rank(volume), 5))"),
    ("alpha014", "-1 * (rank(delta(returns, 3)) * fill(correlation(open, volume, 10)))"),
    ("alpha015", "-1 * ts_sum(rank(fill(correlation_twopass(rank(high), rank(volume), 3))), 3)"),
    ("alpha016", "-1 * rank(covariance(rank(high), rank(volume), 5))"),
    ("alpha017", "-1 * (rank(ts_rank(close, 10)) * rank(delta(delta(close, 1), 1))"
                 " * rank(ts_rank(fill(volume / adv(20)), 5)))"),
    ("alpha018", "-1 * rank((stddev(abs(close - open), 5) + (close - open)) + fill(correlation(close, open, 10)))"),
    ("alpha019", "(-1 * sign((close - delay(close, 7)) + delta(close, 7))) * (1 + rank(1 + ts_sum(returns, 250)))"),
    ("alpha020", "-1 * (rank(open - delay(high, 1)) * rank(open - delay(close, 1)) * rank(open - delay(low, 1)))"),
    ("alpha021", "(~(sma(close, 8) + stddev(close, 8) < sma(close, 2))"
                 " & ((sma(close, 2) < sma(close, 8) - stddev(close, 8)) | (volume / adv(20) >= 1))) ? 1 : -1"),
    ("alpha022", "-1 * delta(fill(correlation(high, volume, 5)), 5) * rank(stddev(close, 20))"),
    ("alpha023", "(sma(high, 20) < high) ? (-1 * fill(delta(high, 2))) : 0"),
    ("alpha024", "(delta(sma(close, 100), 100) / delay(close, 100) <= 0.05) ? (-1 * (close - ts_min(close, 100)))"
                 " : (-1 * delta(close, 3))"),
    ("alpha025", "rank((((-1 * returns) * adv(20)) * vwap) * (high - close))"),
    ("alpha026", "-1 * ts_max(fill(correlation_twopass(ts_rank(volume, 5), ts_rank(high, 5), 5)), 3)"),
    ("alpha027", "(rank(sma_twopass(correlation(rank(volume), rank(vwap), 6), 2) / 2.0) > 0.5) ? -1 : 1"),
    ("alpha028", "scale((fill(correlation(adv(20), low, 5)) + ((high + low) / 2)) - close)"),
    ("alpha029", "ts_min(rank(rank(scale(log(ts_sum(rank(rank(-1 * rank(delta(close - 1, 5)))), 2))))), 5)"
                 " + ts_rank(delay(-1 * returns, 6), 5)"),
    ("alpha030", "(1.0 - rank(sign(delta(close, 1)) + sign(delay(delta(close, 1), 1))"
                 " + sign(delay(delta(close, 1), 2)))) * fill(ts_sum(volume, 5) / ts_sum(volume, 20), 1)"),
    ("alpha031", "rank(rank(rank(decay_linear(-1 * rank(rank(delta(close, 10))), 10)))) + rank(-1 * delta(close, 3))"
                 " + sign(scale(fill(correlation(adv(20), low, 12))))"),
    ("alpha032", "scale((sma(close, 7) / 7) - close) + (20 * scale(correlation(vwap, delay(close, 5), 230)))"),
    ("alpha033", "rank((open / close) - 1)"),
    ("alpha034", "rank(2 - rank(fill(stddev(returns, 2) / stddev(returns, 5), 1.0)) - rank(delta(close, 1)))"),
    ("alpha035", "(ts_rank(volume, 32) * (1 - ts_rank(close + high - low, 16))) * (1 - ts_rank(returns, 32))"),
    ("alpha036", "((((2.21 * rank(correlation(close - open, delay(volume, 1), 15))) + (0.7 * rank(open - close)))"
                 " + (0.73 * rank(ts_rank(delay(-1 * returns, 6), 5)))) + rank(abs(correlation(vwap, adv(20), 6))))"
                 " + (0.6 * rank(((sma(close, 200) / 200) - open) * (close - open)))"),
    ("alpha037", "rank(correlation(delay(open - close, 1), close, 200)) + rank(open - close)"),
    ("alpha038", "-1 * rank(ts_rank(open, 10)) * rank(fill(close / open, 1.0))"),
    ("alpha039", "(-1 * rank(delta(close, 7) * (1 - rank(decay_linear(volume / adv(20), 9)))))"
                 " * (1 + rank(sma(returns, 250)))"),
    ("alpha040", "-1 * rank(stddev(high, 10)) * fill(correlation(high, volume, 10), 1.0)"),
    ("alpha041", "pow(high * low, 0.5) - vwap"),
    ("alpha042", "rank(vwap - close) / rank(vwap + close)"),
    ("alpha043", "ts_rank(fill(volume / adv(20)), 20) * ts_rank(-1 * delta(close, 7), 8)"),
    ("alpha044", "-1 * fill(correlation(high, rank(volume), 5))"),
    ("alpha045", "-1 * (rank(sma(delay(close, 5), 20)) * fill(correlation(close, volume, 2))"
                 " * rank(fill(correlation(ts_sum(close, 5), ts_sum(close, 20), 2), 1)))"),
    ("alpha046", "(((delay(close, 20) - delay(close, 10)) / 10) - ((delay(close, 10) - close) / 10) > 0.25) ? -1"
                 " : ((((delay(close, 20) - delay(close, 10)) / 10) - ((delay(close, 10) - close) / 10) < 0) ? 1"
                 " : (-1 * delta(close, 1)))"),
    ("alpha047", "fill(((rank(1 / close) * volume / adv(20)) * ((high * rank(high - close)) / (sma(high, 5) / 5)))"
                 " - rank(vwap - delay(vwap, 5)))"),
    ("alpha049", "(((delay(close, 20) - delay(close, 10)) / 10) - ((delay(close, 10) - close) / 10) < -0.1) ? 1"
                 " : (-1 * delta(close, 1))"),
    ("alpha050", "-1 * ts_max(rank(fill(correlation(rank(volume), rank(vwap), 5))), 5)"),
    ("alpha051", "(((delay(close, 20) - delay(close, 10)) / 10) - ((delay(close, 10) - close) / 10) < -0.05) ? 1"
                 " : (-1 * delta(close, 1))"),
    ("alpha052", "((-1 * delta(ts_min(low, 5), 5)) * rank((ts_sum(returns, 240) - ts_sum(returns, 20)) / 220))"
                 " * ts_rank(volume, 5)"),
    ("alpha053", "-1 * delta(((close - low) - (high - close)) / nonzero(close - low), 9)"),
    ("alpha054", "-1 * (low - close) * pow(open, 5) / (nonzero(low - high) * pow(close, 5))"),
    ("alpha055", "-1 * fill(correlation(rank((close - ts_min(low, 12)) / nonzero(ts_max(high, 12) - ts_min(low, 12))),"
                 " rank(volume), 6))"),
    ("alpha057", "0 - (1 * ((close - vwap) / decay_linear(rank(ts_argmax(close, 30)), 2)))"),
    ("alpha060", "0 - ((2 * scale(rank(((close - low) - (high - close)) * volume / nonzero(high - low))))"
                 " - scale(rank(ts_argmax(close, 10))))"),
    ("alpha061", "b10(rank(vwap - ts_min(vwap, 16)) < rank(correlation(vwap, adv(180), 18)))"),
    ("alpha062", "b10(rank(correlation(vwap, sma(adv(20), 22), 10))"
                 " < rank(b10((rank(open) + rank(open)) < (rank((high + low) / 2) + rank(high))))) * -1"),
    ("alpha064", "b10(rank(correlation(sma((open * 0.178404) + (low * (1 - 0.178404)), 13), sma(adv(120), 13), 17))"
                 " < rank(delta((((high + low) / 2) * 0.178404) + (vwap * (1 - 0.178404)), 4))) * -1"),
    ("alpha065", "b10(rank(correlation((open * 0.00817205) + (vwap * (1 - 0.00817205)), sma(adv(60), 9), 6))"
                 " < rank(open - ts_min(open, 14))) * -1"),
    ("alpha066", "(rank(decay_linear(delta(vwap, 4), 7)) + ts_rank(decay_linear(fill((((low * 0.96633)"
                 " + (low * (1 - 0.96633))) - vwap) / (open - ((high + low) / 2))), 11), 7)) * -1"),
    ("alpha068", "b10(ts_rank(correlation(rank(high), rank(adv(15)), 9), 14)"
                 " < rank(delta((close * 0.518371) + (low * (1 - 0.518371)), 1))) * -1"),
    ("alpha071", "max(ts_rank(decay_linear(correlation(ts_rank(close, 3), ts_rank(adv(180), 12), 18), 4), 16),"
                 " ts_rank(decay_linear(pow(rank((low + open) - (vwap + vwap)), 2), 16), 4))"),
    ("alpha072", "(rank(decay_linear(correlation((high + low) / 2, adv(40), 9), 10)) + 0.0001)"
                 " / (rank(decay_linear(correlation(ts_rank(vwap, 4), ts_rank(volume, 19), 7), 3)) + 0.0001)"),
    ("alpha073", "-1 * max(rank(decay_linear(delta(vwap, 5), 3)), ts_rank(decay_linear((delta((open * 0.147155)"
                 " + (low * (1 - 0.147155)), 2) / ((open * 0.147155) + (low * (1 - 0.147155)))) * -1, 3), 17))"),
    ("alpha074", "b10(rank(correlation(close, sma(adv(30), 37), 15)) < rank(correlation_twopass(rank((high"
                 " * 0.0261661) + (vwap * (1 - 0.0261661))), rank(volume), 11))) * -1"),
    ("alpha075", "b10(rank(correlation(vwap, volume, 4)) < rank(correlation(rank(low), rank(adv(50)), 12)))"),
    ("alpha077", "min(rank(decay_linear((((high + low) / 2) + high) - (vwap + high), 20)),"
                 " rank(decay_linear(fill(correlation((high + low) / 2, adv(40), 3), 1), 6)))"),
    ("alpha078", "pow(rank(correlation(ts_sum((low * 0.352233) + (vwap * (1 - 0.352233)), 20), ts_sum(adv(40), 20),"
                 " 7)), rank(fill(correlation(rank(vwap), rank(volume), 6))))"),
    ("alpha081", "b10(rank(log(product(rank(pow(rank(correlation(vwap, ts_sum(adv(10), 50), 8)), 4)), 15)))"
                 " < rank(correlation(rank(vwap), rank(volume), 5))) * -1"),
    ("alpha083", "(rank(delay((high - low) / (ts_sum(close, 5) / 5), 2)) * rank(rank(volume)))"
                 " / (((high - low) / (ts_sum(close, 5) / 5)) / (vwap - close))"),
    ("alpha084", "pow(ts_rank(vwap - ts_max(vwap, 15), 21), delta(close, 5))"),
    ("alpha085", "pow(rank(fill(correlation((high * 0.876703) + (close * (1 - 0.876703)), adv(30), 10), 1)),"
                 " rank(fill(correlation(ts_rank((high + low) / 2, 4), ts_rank(volume, 10), 7), 1)))"),
    ("alpha086", "b10(ts_rank(correlation(close, sma(adv(20), 15), 6), 20)"
                 " < rank((open + close) - (vwap + open))) * -1"),
    ("alpha088", "min(ts_rank(decay_linear(fill(correlation(ts_rank(close, 8), ts_rank(adv(60), 21), 8)), 7), 3),"
                 " rank(decay_linear((rank(open) + rank(low)) - (rank(high) + rank(close)), 8)))"),
    ("alpha092", "min(ts_rank(decay_linear(fill(correlation(rank(low), rank(adv(30)), 8)), 7), 7),"
                 " ts_rank(decay_linear(b10((((high + low) / 2) + close) < (low + open)), 15), 19))"),
    ("alpha094", "pow(rank(vwap - ts_min(vwap, 12)),"
                 " ts_rank(fill(correlation(ts_rank(vwap, 20), ts_rank(adv(60), 4), 18)), 3)) * -1"),
    ("alpha095", "b10(rank(open - ts_min(open, 12)) < ts_rank(pow(rank(correlation(sma((high + low) / 2, 19),"
                 " sma(adv(40), 19), 13)), 5), 12))"),
    ("alpha096", "-1 * max(ts_rank(decay_linear(fill(correlation(rank(vwap), rank(volume), 4)), 4), 8),"
                 " ts_rank(decay_linear(ts_argmax(fill(correlation(ts_rank(close, 7), ts_rank(adv(60), 4), 4)), 13),"
                 " 14), 13))"),
    ("alpha098", "rank(decay_linear(correlation(vwap, sma(adv(5), 26), 5), 7))"
                 " - rank(decay_linear(ts_rank(ts_argmin(fill(correlation(rank(open), rank(adv(15)), 21)), 9), 7), 8))"),
    ("alpha099", "b10(rank(correlation(ts_sum((high + low) / 2, 20), ts_sum(adv(60), 20), 9))"
                 " < rank(correlation(low, volume, 6))) * -1"),
    ("alpha101", "(close - open) / ((high - low) + 0.001)"),
]


# ---- functions of the language -------------------------------------------
def _op(x):
    return ConstantOp(x) if isinstance(x, (int, float)) else x


def _corr(a, b, n, two_pass=False):
    r = WindowedCorrelation(a, int(n), b)
    if two_pass:
        r.attrs["no_fast_stat"] = True
    return r


def _sma(x, n, two_pass=False):
    r = WindowedAvg(x, int(n))
    if two_pass:
        r.attrs["no_fast_stat"] = True
    return r


_FUNCS: Dict[str, Callable] = {
    "rank": Rank,
    "scale": Scale,
    "delay": lambda x, n: BackRef(x, int(n)),
    "delta": lambda x, n=1: x - BackRef(x, int(n)),
    "correlation": _corr,
    "correlation_twopass": lambda a, b, n: _corr(a, b, n, True),
    "covariance": lambda a, b, n: WindowedCovariance(a, int(n), b),
    "stddev": lambda x, n: WindowedStddev(x, int(n)),
    "sma": _sma,
    "sma_twopass": lambda x, n: _sma(x, n, True),
    "ts_sum": lambda x, n: WindowedSum(x, int(n)),
    "ts_min": lambda x, n: WindowedMin(x, int(n)),
    "ts_max": lambda x, n: WindowedMax(x, int(n)),
    "ts_argmax": lambda x, n: TsArgMax(x, int(n)),
    "ts_argmin": lambda x, n: TsArgMin(x, int(n)),
    "ts_rank": lambda x, n: TsRank(x, int(n)),
    "decay_linear": lambda x, n: DecayLinear(x, int(n)),
    "product": lambda x, n: WindowedProduct(x, int(n)),
    "sign": Sign,
    "abs": Abs,
    "log": Log,
    "pow": lambda a, b: Pow(a, _op(b)),
    "max": lambda a, b: Max(_op(a), _op(b)),
    "min": lambda a, b: Min(_op(a), _op(b)),
    # inf and NaN replaced by v
    "fill": lambda x, v=0.0: SetInfOrNanToValue(x, v),
    # a zero divisor replaced by 0.0001
    "nonzero": lambda x: Select(Equals(x, ConstantOp(0)), ConstantOp(0.0001), x),
    # a condition as 1 / 0
    "b10": lambda c: Select(c, ConstantOp(1), ConstantOp(0)),
    "sq": lambda x: x * x,
}

# ---- parser (precedence climbing) ----------------------------------------
_TOKEN = re.compile(r"\s*(?:(\d+\.\d*|\d*\.\d+|\d+)|([A-Za-z_]\w*)|(<=|>=|==|[-+*/<>()?:,&|~]))")

# binary operators: precedence, then how two values combine
_BINARY = {
    "|": (1, lambda a, b: a | b), "&": (2, lambda a, b: a & b),
    "<": (3, lambda a, b: a < b), "<=": (3, lambda a, b: a <= b), ">": (3, lambda a, b: a > b),
    ">=": (3, lambda a, b: a >= b),
    "==": (3, lambda a, b: (a == b) if _num(a) and _num(b) else Equals(_op(a), _op(b))),
    "+": (4, lambda a, b: a + b), "-": (4, lambda a, b: a - b),
    "*": (5, lambda a, b: a * b), "/": (5, lambda a, b: a / b),
}


def _num(x):
    return isinstance(x, (int, float))


def _tokens(text: str):
    pos, out = 0, []
    text = text.rstrip()
    while pos < len(text):
        m = _TOKEN.match(text, pos)
        if not m:
            raise SyntaxError(f"bad formula text at {text[pos:pos + 20]!r}")
        num, name, sym = m.groups()
        if num is not None:
            out.append(("num", float(num) if ("." in num) else int(num)))
        elif name is not None:
            out.append(("name", name))
        else:
            out.append(("sym", sym))
        pos = m.end()
    out.append(("end", None))
    return out


class _Parser:
    def __init__(self, text: str, env: Dict[str, object]):
        self.toks = _tokens(text)
        self.i = 0
        self.env = env

    def peek(self):
        return self.toks[self.i]

    def take(self, sym=None):
        t = self.toks[self.i]
        if sym is not None and t != ("sym", sym):
            raise SyntaxError(f"expected {sym!r}, found {t[1]!r}")
        self.i += 1
        return t

    def parse(self):
        v = self.ternary()
        if self.peek()[0] != "end":
            raise SyntaxError(f"trailing text at {self.peek()[1]!r}")
        return v

    def ternary(self):
        cond = self.binary(1)
        if self.peek() == ("sym", "?"):
            self.take("?")
            a = self.ternary()
            self.take(":")
            b = self.ternary()
            return Select(cond, _op(a), _op(b))
        return cond

    def binary(self, level):
        lhs = self.unary()
        while True:
            t = self.peek()
            if t[0] != "sym" or t[1] not in _BINARY or _BINARY[t[1]][0] < level:
                return lhs
            prec, fn = _BINARY[t[1]]
            self.take()
            rhs = self.binary(prec + 1)
            lhs = fn(lhs, rhs)

    def unary(self):
        t = self.peek()
        if t == ("sym", "-"):
            self.take()
            v = self.unary()
            if not _num(v):
                raise SyntaxError("unary minus applies to numbers only (write -1 * x)")
            return -v
        if t == ("sym", "~"):
            self.take()
            return ~self.unary()
        return self.primary()

    def primary(self):
        kind, val = self.take()
        if kind == "num":
            return val
        if kind == "sym" and val == "(":
            v = self.ternary()
            self.take(")")
            return v
        if kind == "name":
            if self.peek() == ("sym", "("):
                self.take("(")
                args = []
                if self.peek() != ("sym", ")"):
                    args.append(self.ternary())
                    while self.peek() == ("sym", ","):
                        self.take(",")
                        args.append(self.ternary())
                self.take(")")
                if val == "adv":
                    return _sma(self.env["volume"], args[0])
                if val not in _FUNCS:
                    raise SyntaxError(f"unknown function {val}")
                return _FUNCS[val](*args)
            if val not in self.env:
                raise SyntaxError(f"unknown name {val}")
            return self.env[val]
        raise SyntaxError(f"unexpected {val!r}")


def build_formula(text: str, d: AllData) -> OpBase:
    env = {k: getattr(d, k) for k in ("open", "close", "high", "low", "volume", "amount", "vwap", "returns")}
    return _op(_Parser(text, env).parse())


def _alpha(name: str, text: str):
    def f(d: AllData) -> OpBase:
        return build_formula(text, d)
    f.__name__ = f.__qualname__ = name
    f.__doc__ = f"{name}: {text}"
    return f


FORMULAS: Dict[str, str] = dict(_FORMULAS)
all_alpha = [_alpha(n, t) for n, t in _FORMULAS]
globals().update({f.__name__: f for f in all_alpha})
