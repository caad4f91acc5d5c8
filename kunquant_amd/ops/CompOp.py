"""Compositive (user-level) ops and their decompositions into primitives.

The decompositions define the arithmetic every backend must reproduce, so
the primitive graphs built here follow KunQuant/ops/CompOp.py op for op
(same primitives, same construction order):
  WindowedReduce 26-37, WindowedSum 39-46, WindowedProduct 48-55,
  min/max with skip-list switch 14-17, 57-102, fast-stat helpers 105-153,
  WindowedAvg 155-179, WindowedVar 181-235, WindowedStddev 237-249,
  WindowedCovariance 251-282, WindowedCorrelation 284-325,
  skew/kurt 327-460, TsArgMax/TsArgMin/TsRank 463-510, Clip 512-529,
  DecayLinear 531-546, Pow 548-617, linear regression 620-654,
  WindowedMaxDrawdown 657-678, WindowedQuantile 681-700.
"""
import math
from typing import List

from ..Op import (Builder, CompositiveOp, ConstantOp, ForeachBackWindow, IterValue, OpBase,
                  WindowLoopIndex, WindowedTempOutput, WindowedTrait)
from .ElewiseOp import (Abs, And, DivConst, Div, Equals, Exp, Log, Max, Min, Mul, Select, Sqrt,
                        Sub)
from .MiscOp import (Accumulator, BackRef, ReturnFirstValue, SetAccumulator, SkipListArgMin,
                     SkipListMax, SkipListMin, SkipListQuantile, SkipListRank, SkipListState,
                     WindowedLinearRegression, WindowedLinearRegressionResiImpl,
                     WindowedLinearRegressionRSqaureImpl, WindowedLinearRegressionSlopeImpl)
from .ReduceOp import (ReduceAdd, ReduceArgMax, ReduceArgMin, ReduceDecayLinear, ReduceMax,
                       ReduceMin, ReduceMul, ReduceRank)


def _is_fast_stat(opt: dict, attrs: dict) -> bool:
    """Rolling O(1) statistics are used unless `no_fast_stat` is set globally
    (default True) or on the op."""
    return not opt.get("no_fast_stat", True) and not attrs.get("no_fast_stat", False)


def _decide_use_skip_list(window: int, blocking_len: int) -> bool:
    """Cost model of the reference: sorted-state ops beat an O(w) scan when
    log2(w) * blocking_len * 5 < w."""
    return math.log2(window) * blocking_len * 5 < window


class WindowedCompositiveOp(CompositiveOp, WindowedTrait):
    def __init__(self, v: OpBase, window: int, v2=None) -> None:
        super().__init__([v] if v2 is None else [v, v2], [("window", window)])


class WindowedReduce(WindowedCompositiveOp):
    """Loop over a temp window of the input and reduce it."""

    def make_reduce(self, v: OpBase, newest: OpBase) -> OpBase:
        raise RuntimeError("Not implemented")

    def decompose(self, options: dict) -> List[OpBase]:
        window = self.attrs["window"]
        b = Builder(self.get_parent())
        with b:
            buf = WindowedTempOutput(self.inputs[0], window)
            loop = ForeachBackWindow(buf, window)
            self.make_reduce(IterValue(loop, buf), self.inputs[0])
        return b.ops


class WindowedSum(WindowedReduce):
    """rolling(window).sum(); NaN for the first window-1 rows."""

    def make_reduce(self, v, newest):
        return ReduceAdd(v)


class WindowedProduct(WindowedReduce):
    """rolling(window).product()."""

    def make_reduce(self, v, newest):
        return ReduceMul(v)


class _WindowedMinMaxBase(WindowedReduce):
    def on_skip_list(self, skplist: SkipListState, cur: OpBase) -> OpBase:
        raise RuntimeError("Not implemented")

    def decompose(self, options: dict) -> List[OpBase]:
        window = self.attrs["window"]
        if not _decide_use_skip_list(window, options["blocking_len"]):
            return super().decompose(options)
        b = Builder(self.get_parent())
        with b:
            cur = self.inputs[0]
            state = SkipListState(BackRef(cur, window), cur, window)
            self.on_skip_list(state, cur)
        return b.ops


class WindowedMin(_WindowedMinMaxBase):
    """rolling(window).min()."""

    def make_reduce(self, v, newest):
        return ReduceMin(v)

    def on_skip_list(self, skplist, cur):
        return SkipListMin(skplist)


class WindowedMax(_WindowedMinMaxBase):
    """rolling(window).max()."""

    def make_reduce(self, v, newest):
        return ReduceMax(v)

    def on_skip_list(self, skplist, cur):
        return SkipListMax(skplist)


# ---------------------------------------------------------------------------
# Fast (rolling add/remove) statistics, pandas-style with Kahan-compensated
# mean updates. Used when no_fast_stat is False.

def _kahan_sub(mask, total, small, comp):
    y = small - comp
    t = y - total
    comp = SetAccumulator(comp, mask, t + total - y)
    t = ReturnFirstValue([t, comp])
    return t, comp


def _avg_remove_window(oldx: OpBase, window: int):
    notnan_old = Equals(oldx, oldx)
    count_acc = Accumulator(notnan_old, f"avg_obs_{window}")
    count = Select(notnan_old, count_acc - 1, count_acc)
    count_nonzero = Abs(count) > 0.001
    can_update = And(count_nonzero, notnan_old)
    mean_acc = Accumulator(oldx, f"avg_{window}")
    mean = mean_acc
    comp_remove = Accumulator(mean, f"avg_comp_remove_{window}")
    delta, _ = _kahan_sub(can_update, mean, oldx, comp_remove)
    mean = Select(can_update, mean - delta / count, mean)
    return count_acc, count, can_update, mean_acc, mean, comp_remove, delta


def _update_accumulator(acc, mask, current_value, update_value):
    return SetAccumulator(acc, mask, Select(mask, update_value, current_value))


def _avg_add_window(x, window, count_acc, count, mean, mean_acc):
    notnan = Equals(x, x)
    count = _update_accumulator(count_acc, notnan, count, count + 1)
    count_nonzero = Abs(count) > 0.001
    comp_add = Accumulator(count_nonzero, f"avg_comp_add_{window}")
    delta, _ = _kahan_sub(notnan, mean, x, comp_add)
    mean = _update_accumulator(mean_acc, notnan, mean,
                               Select(count_nonzero, mean + delta / count, ConstantOp(0)))
    return count, notnan, mean, comp_add, delta


def with_same_opt_flag(target: OpBase, current: dict) -> OpBase:
    if "no_fast_stat" in current:
        target.attrs["no_fast_stat"] = current["no_fast_stat"]
    return target


class WindowedAvg(WindowedCompositiveOp):
    """rolling(window).mean()."""

    def decompose(self, options: dict) -> List[OpBase]:
        window = self.attrs["window"]
        x = self.inputs[0]
        b = Builder(self.get_parent())
        with b:
            if _is_fast_stat(options, self.attrs):
                oldx = BackRef(x, window)
                count_acc, count, _, mean_acc, mean, _, _ = _avg_remove_window(oldx, window)
                count, _, mean, _, _ = _avg_add_window(x, window, count_acc, count, mean, mean_acc)
                Select(Abs(count - window) < 0.001, mean, ConstantOp("nan"))
            else:
                DivConst(WindowedSum(x, window), window)
        return b.ops


def _stddev_remove_window(oldx, window):
    count_acc, count, can_update, mean_acc, mean, comp_remove, _ = _avg_remove_window(oldx, window)
    prev_mean = mean_acc - comp_remove
    var_acc = Accumulator(mean, f"var_{window}")
    var = Select(can_update, var_acc - (oldx - prev_mean) * (oldx - mean), var_acc)
    return count_acc, count, can_update, mean_acc, mean, comp_remove, var_acc, var


def _stddev_add_window(x, window, count_acc, count, oldmean, mean_acc, var, var_acc):
    count, notnan, new_mean, comp_add, _ = _avg_add_window(x, window, count_acc, count, oldmean, mean_acc)
    prev_mean = oldmean - comp_add
    var = _update_accumulator(var_acc, notnan, var, var + (x - prev_mean) * (x - new_mean))
    return count, notnan, new_mean, var


class WindowedVar(WindowedCompositiveOp):
    """rolling(window).var() (unbiased)."""

    def decompose(self, options: dict) -> List[OpBase]:
        window = self.attrs["window"]
        b = Builder(self.get_parent())
        with b:
            if _is_fast_stat(options, self.attrs):
                x = self.inputs[0]
                oldx = BackRef(x, window)
                count_acc, count, _, mean_acc, mean, _, var_acc, var = _stddev_remove_window(oldx, window)
                count, _, mean, var = _stddev_add_window(x, window, count_acc, count, mean, mean_acc,
                                                         var, var_acc)
                var = DivConst(var, window - 1)
                Select(Abs(count - window) < 0.001, var, ConstantOp("nan"))
            else:
                avg = with_same_opt_flag(WindowedAvg(self.inputs[0], window), self.attrs)
                buf = WindowedTempOutput(self.inputs[0], window)
                loop = ForeachBackWindow(buf, window)
                b.set_loop(loop)
                diff = Sub(IterValue(loop, buf), avg)
                sqr = Mul(diff, diff)
                b.set_loop(self.get_parent())
                DivConst(ReduceAdd(sqr), window - 1)
        return b.ops


class WindowedStddev(WindowedCompositiveOp):
    """rolling(window).std() (unbiased)."""

    def decompose(self, options: dict) -> List[OpBase]:
        b = Builder(self.get_parent())
        with b:
            Sqrt(with_same_opt_flag(WindowedVar(self.inputs[0], self.attrs["window"]), self.attrs))
        return b.ops


class WindowedCovariance(WindowedCompositiveOp):
    """rolling(window).cov(y)."""

    def decompose(self, options: dict) -> List[OpBase]:
        window = self.attrs["window"]
        x, y = self.inputs
        b = Builder(self.get_parent())
        with b:
            if _is_fast_stat(options, self.attrs):
                mxy = WindowedAvg(x * y, window)
                mx = WindowedAvg(x, window)
                my = WindowedAvg(y, window)
                (mxy - mx * my) * (window / (window - 1))
            else:
                ax = with_same_opt_flag(WindowedAvg(x, window), self.attrs)
                ay = with_same_opt_flag(WindowedAvg(y, window), self.attrs)
                wx = WindowedTempOutput(x, window)
                wy = WindowedTempOutput(y, window)
                loop = ForeachBackWindow(wx, window, wy)
                b.set_loop(loop)
                dx = Sub(IterValue(loop, wx), ax)
                dy = Sub(IterValue(loop, wy), ay)
                prod = Mul(dx, dy)
                b.set_loop(self.get_parent())
                ReduceAdd(prod) / (window - 1)
        return b.ops


class WindowedCorrelation(WindowedCompositiveOp):
    """rolling(window).corr(y)."""

    def decompose(self, options: dict) -> List[OpBase]:
        window = self.attrs["window"]
        x, y = self.inputs
        b = Builder(self.get_parent())
        with b:
            if _is_fast_stat(options, self.attrs):
                mxy = WindowedAvg(x * y, window)
                mx = WindowedAvg(x, window)
                my = WindowedAvg(y, window)
                vx = WindowedVar(x, window)
                vy = WindowedVar(y, window)
                num = (mxy - mx * my) * (window / (window - 1))
                num / Sqrt(vx * vy)
            else:
                ax = with_same_opt_flag(WindowedAvg(x, window), self.attrs)
                ay = with_same_opt_flag(WindowedAvg(y, window), self.attrs)
                wx = WindowedTempOutput(x, window)
                wy = WindowedTempOutput(y, window)
                loop = ForeachBackWindow(wx, window, wy)
                b.set_loop(loop)
                dx = Sub(IterValue(loop, wx), ax)
                dy = Sub(IterValue(loop, wy), ay)
                sx = Mul(dx, dx)
                sy = Mul(dy, dy)
                sxy = Mul(dx, dy)
                b.set_loop(self.get_parent())
                cross = ReduceAdd(sxy)
                nx = Sqrt(ReduceAdd(sx))
                ny = Sqrt(ReduceAdd(sy))
                Div(cross, Mul(nx, ny))
        return b.ops


def _make_m2_m3_acc_skew(oldx, window):
    return Accumulator(oldx, f"skew_m2_{window}"), Accumulator(oldx, f"skew_m3_{window}")


def _remove_skew_update(m2_acc, m3_acc, notnan_old, delta_mean, n):
    delta_n = delta_mean / n
    term1 = delta_n * delta_mean * (n + 1.0)
    m3_update = delta_n * (term1 * (n + 2.0) - 3.0 * m2_acc)
    m3 = Select(notnan_old, m3_acc - m3_update, m3_acc)
    m2 = Select(notnan_old, m2_acc - term1, m2_acc)
    return m2, m3, delta_n, term1


def _add_skew_update(m2_acc, m3_acc, notnan, delta_mean, n, m2, m3):
    delta_n = delta_mean / n
    term1 = delta_mean * delta_n * (n - 1.0)
    m3_update = delta_n * (term1 * (n - 2.0) - 3.0 * m2)
    m3 = _update_accumulator(m3_acc, notnan, m3, m3 + m3_update)
    m2 = _update_accumulator(m2_acc, notnan, m2, m2 + term1)
    return m2, m3, delta_n, term1


class WindowedSkew(WindowedCompositiveOp):
    """rolling(window).skew()."""

    def decompose(self, options: dict) -> List[OpBase]:
        window = self.attrs["window"]
        x = self.inputs[0]
        b = Builder(self.get_parent())
        with b:
            if _is_fast_stat(options, self.attrs):
                oldx = BackRef(x, window)
                count_acc, n, notnan_old, mean_acc, mean, _, delta_mean = _avg_remove_window(oldx, window)
                m2_acc, m3_acc = _make_m2_m3_acc_skew(oldx, window)
                m2, m3, _, _ = _remove_skew_update(m2_acc, m3_acc, notnan_old, delta_mean, n)
                n, notnan, mean, _, delta_mean = _avg_add_window(x, window, count_acc, n, mean, mean_acc)
                m2, m3, _, _ = _add_skew_update(m2_acc, m3_acc, notnan, delta_mean, n, m2, m3)
                ratio = m3 / (m2 * Sqrt(m2))
                result = ratio * (window * math.sqrt(window - 1) / (window - 2))
                result = ReturnFirstValue([result, mean])
                Select(And(Abs(n - window) < 0.001, m2 > n * 1e-14), result, ConstantOp("nan"))
            else:
                avg = WindowedAvg(x, window)
                buf = WindowedTempOutput(x, window)
                loop = ForeachBackWindow(buf, window)
                b.set_loop(loop)
                d = Sub(IterValue(loop, buf), avg)
                d2 = Mul(d, d)
                d3 = d2 * d
                b.set_loop(self.get_parent())
                s2 = ReduceAdd(d2)
                s3 = ReduceAdd(d3)
                s3 / (s2 * Sqrt(s2)) * ConstantOp(math.sqrt(window - 1) * window / (window - 2))
        return b.ops


class WindowedKurt(WindowedCompositiveOp):
    """rolling(window).kurt()."""

    def decompose(self, options: dict) -> List[OpBase]:
        n = self.attrs["window"]
        x = self.inputs[0]
        b = Builder(self.get_parent())
        with b:
            if _is_fast_stat(options, self.attrs):
                window = self.attrs["window"]
                oldx = BackRef(x, window)
                count_acc, n, notnan_old, mean_acc, mean, _, delta_mean = _avg_remove_window(oldx, window)
                m2_acc, m3_acc = _make_m2_m3_acc_skew(oldx, window)
                m4_acc = Accumulator(oldx, f"kurt_m4_{window}")
                m2, m3, delta_n, term1 = _remove_skew_update(m2_acc, m3_acc, notnan_old, delta_mean, n)
                m4_update = delta_n * (4 * m3_acc + delta_n * (6 * m2_acc - term1 * (n * n + 3 * n + 3)))
                m4 = Select(notnan_old, m4_acc + m4_update, m4_acc)
                n, notnan, mean, _, delta_mean = _avg_add_window(x, window, count_acc, n, mean, mean_acc)
                old_m2, old_m3 = m2, m3
                m2, m3, delta_n, term1 = _add_skew_update(m2_acc, m3_acc, notnan, delta_mean, n, m2, m3)
                m4_update = delta_n * (-4 * old_m3 + delta_n * (6 * old_m2 + term1 * (n * n - 3 * n + 3)))
                m4 = _update_accumulator(m4_acc, notnan, m4, m4 + m4_update)
                ratio = m4 / (m2 * m2)
                inner = (window * (window + 1)) * ratio - 3 * (window - 1)
                result = ((window - 1) / ((window - 2) * (window - 3))) * inner
                result = ReturnFirstValue([result, mean, m3])
                Select(And(Abs(n - window) < 0.001, m2 > n * 1e-14), result, ConstantOp("nan"))
            else:
                avg = WindowedAvg(x, n)
                buf = WindowedTempOutput(x, n)
                with ForeachBackWindow(buf, n) as loop:
                    d = Sub(IterValue(loop, buf), avg)
                    d2 = Mul(d, d)
                    d4 = d2 * d2
                s2 = ReduceAdd(d2)
                s4 = ReduceAdd(d4)
                factor = n * (n - 1) * (n + 1) / (n - 2) / (n - 3)
                offset = 3 * (n - 1) * (n - 1) / (n - 2) / (n - 3)
                s4 / (s2 * s2) * factor - offset
        return b.ops


class TsArgMax(WindowedReduce):
    """rolling(window).apply(np.argmax) + 1 (oldest row is 1)."""

    def decompose(self, options: dict) -> List[OpBase]:
        window = self.attrs["window"]
        if _decide_use_skip_list(window, options["blocking_len"]):
            b = Builder(self.get_parent())
            with b:
                TsArgMin(0 - self.inputs[0], window)
            return b.ops
        return super().decompose(options)

    def make_reduce(self, v, newest):
        return self.attrs["window"] - ReduceArgMax(v)


class TsArgMin(_WindowedMinMaxBase):
    """rolling(window).apply(np.argmin) + 1 (oldest row is 1)."""

    def on_skip_list(self, skplist, cur):
        return SkipListArgMin([skplist], [])

    def make_reduce(self, v, newest):
        return self.attrs["window"] - ReduceArgMin(v)


class TsRank(_WindowedMinMaxBase):
    """rolling(window).rank() of the newest value."""

    def on_skip_list(self, skplist, cur):
        return SkipListRank(skplist, cur)

    def make_reduce(self, v, newest):
        return ReduceRank(v, newest)


class Clip(CompositiveOp):
    """max(min(v, eps), -eps)."""

    def __init__(self, v: OpBase, eps: float) -> None:
        super().__init__([v], [("value", eps)])

    def decompose(self, options: dict) -> List[OpBase]:
        eps = self.attrs["value"]
        b = Builder(self.get_parent())
        with b:
            Max(Min(self.inputs[0], ConstantOp(eps)), ConstantOp(-eps))
        return b.ops


class DecayLinear(WindowedCompositiveOp):
    """Linearly decaying weighted average; weight of the k-th oldest row is
    k / ((1 + w) * w / 2)."""

    def decompose(self, options: dict) -> List[OpBase]:
        window = self.attrs["window"]
        b = Builder(self.get_parent())
        with b:
            buf = WindowedTempOutput(self.inputs[0], window)
            loop = ForeachBackWindow(buf, window)
            ReduceDecayLinear(IterValue(loop, buf), None, [("window", window)])
        return b.ops


def _pow_by_constant(base, expo):
    """Exact-structure lowering of base ** expo for special constant
    exponents; None when a generic exp(log(base) * expo) is needed."""
    if expo == 0:
        return ConstantOp(1)
    if expo < 0:
        positive = _pow_by_constant(base, -expo)
        return ConstantOp(1) / positive
    if expo == 0.5:
        return Sqrt(base)
    if abs(expo - int(expo) - 0.5) < 1e-5:
        intpart = _pow_by_constant(base, int(expo))
        if intpart is None:
            return None
        return Sqrt(base) * intpart
    if int(expo) == expo and 0 <= expo <= 1024:
        # binary exponentiation: square the base, multiply set bits in
        remaining = int(expo)
        square = base
        bit = 1
        acc = base if remaining & bit else 1
        remaining &= ~bit
        while remaining:
            bit *= 2
            square = square * square
            if remaining & bit:
                acc = acc * square
            remaining &= ~bit
        return acc
    return None


class Pow(CompositiveOp):
    """base ** expo."""

    def __init__(self, base: OpBase, expo: OpBase) -> None:
        super().__init__([base, expo], None)

    def decompose(self, options: dict) -> List[OpBase]:
        b = Builder(self.get_parent())
        base, expo = self.inputs
        if isinstance(base, ConstantOp):
            ln_base = math.log(base.attrs["value"])
            with b:
                if abs(ln_base) < 1e-5:
                    ConstantOp(1)
                elif abs(ln_base - 1) < 1e-5:
                    Exp(expo)
                else:
                    Exp(expo * ln_base)
            return b.ops
        if isinstance(expo, ConstantOp):
            ev = expo.attrs["value"]
            with b:
                if _pow_by_constant(base, ev) is None:
                    Exp(Log(base) * ev)
            return b.ops
        with b:
            Exp(expo * Log(base))
        return b.ops


class WindowedLinearRegressionBase(WindowedCompositiveOp):
    def make_extract(self, regr: OpBase, inp: OpBase) -> OpBase:
        raise RuntimeError("Not implemented")

    def decompose(self, options: dict) -> List[OpBase]:
        b = Builder(self.get_parent())
        with b:
            regr = WindowedLinearRegression(self.inputs[0], self.attrs["window"])
            self.make_extract(regr, self.inputs[0])
        return b.ops


class WindowedLinearRegressionRSqaure(WindowedLinearRegressionBase):
    """R-squared of the rolling regression against 1..window."""

    def make_extract(self, regr, inp):
        return WindowedLinearRegressionRSqaureImpl(regr)


class WindowedLinearRegressionSlope(WindowedLinearRegressionBase):
    """Slope of the rolling regression against 1..window."""

    def make_extract(self, regr, inp):
        return WindowedLinearRegressionSlopeImpl(regr)


class WindowedLinearRegressionResi(WindowedLinearRegressionBase):
    """Residual of the newest value against the rolling regression."""

    def make_extract(self, regr, inp):
        return WindowedLinearRegressionResiImpl(regr, inp)


class WindowedMaxDrawdown(WindowedCompositiveOp):
    """(max - trough) / max over the window, trough taken at or after the
    (oldest) maximum."""

    def decompose(self, options: dict) -> List[OpBase]:
        window = self.attrs["window"]
        v = self.inputs[0]
        b = Builder(self.get_parent())
        with b:
            peak = WindowedMax(v, window)
            peak_idx = TsArgMax(v, window) - 1
            inf = ConstantOp(float("inf"))
            with ForeachBackWindow(v, window) as loop:
                idx = WindowLoopIndex(loop)
                kept = Select(idx >= peak_idx, IterValue(loop, v), inf)
            trough = ReduceMin(kept)
            (peak - trough) / peak
        return b.ops


class WindowedQuantile(CompositiveOp, WindowedTrait):
    """rolling(window).quantile(q, interpolation='linear')."""

    def __init__(self, v: OpBase, window: int, q: float) -> None:
        super().__init__([v], [("window", window), ("q", q)])

    def required_input_window(self) -> int:
        return self.attrs["window"] + 1

    def decompose(self, options: dict) -> List[OpBase]:
        window = self.attrs["window"]
        v = self.inputs[0]
        b = Builder(self.get_parent())
        with b:
            state = SkipListState(BackRef(v, window), v, window)
            SkipListQuantile(state, self.attrs["q"])
        return b.ops
