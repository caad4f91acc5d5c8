"""Algebraic rewrites and the FastWindowedSum rewrite
(KunQuant/passes/SpecialOpt.py:17-240).

Rewrites, tried in this order on every op until a fixed point:
  x*1, 1*x -> x ; x*(-1) -> 0-x
  log(a) - log(b) -> log(a/b), also through BackRef of a log window
  Rank(T(x)) -> Rank(x) for increasing T
  a/b <cmp> c -> a <cmp> b*c (and the mirrored form)
  Sign(T(x)) -> Sign(x) for sign-preserving T
  Abs(x) -> x when x is known non-negative
  top-level ReduceAdd over a window loop -> FastWindowedSum (opt_reduce)
They change floating-point results, so every backend must apply them."""
import math
from dataclasses import dataclass
from typing import Dict, List

from ..Op import (Builder, ConstantOp, ForeachBackWindow, IterValue, OpBase, Rank, Scale,
                  WindowedTempOutput)
from ..ops import (Abs, AddConst, BackRef, CmpOp, Div, DivConst, Equals, FastWindowedSum, Log,
                   Mul, MulConst, Pow, ReduceAdd, Sign, Sqrt, Sub, SubConst, TsRank)
from .Util import kun_pass


@dataclass
class _Range:
    start: float
    end: float
    include_start: bool
    include_end: bool


class _Ranges:
    def __init__(self) -> None:
        self.memo: Dict[OpBase, _Range] = {}

    def get(self, op: OpBase) -> _Range:
        r = self.memo.get(op)
        if r is None:
            if isinstance(op, Rank):
                r = _Range(0.0, 1.0, True, True)
            elif isinstance(op, TsRank):
                r = _Range(0.0, op.attrs["window"] - 1, True, True)
            elif isinstance(op, ConstantOp):
                v = op.attrs["value"]
                r = _Range(v, v, True, True)
            else:
                r = _Range(-math.inf, math.inf, False, False)
            self.memo[op] = r
        return r


def _is_one(op: OpBase) -> bool:
    return isinstance(op, ConstantOp) and op.attrs["value"] == 1


def _mul_one(ranges, op):
    if isinstance(op, MulConst):
        if op.attrs["value"] == -1:
            return SubConst(op.inputs[0], 0, True)
        if op.attrs["value"] == 1:
            return op.inputs[0]
    if isinstance(op, Mul):
        if _is_one(op.inputs[0]):
            return op.inputs[1]
        if _is_one(op.inputs[1]):
            return op.inputs[0]
    return None


def _log_difference(ranges, op):
    if not isinstance(op, Sub):
        return None
    lhs, rhs = op.inputs
    if not isinstance(lhs, Log):
        return None
    if isinstance(rhs, Log):
        return Log(Div(lhs.inputs[0], rhs.inputs[0]))
    if (isinstance(rhs, BackRef) and isinstance(rhs.inputs[0], WindowedTempOutput)
            and isinstance(rhs.inputs[0].inputs[0], Log)):
        wbuf = rhs.inputs[0]
        inner = wbuf.inputs[0].inputs[0]
        shifted = BackRef(WindowedTempOutput(inner, wbuf.attrs["window"]), rhs.attrs["window"])
        return Log(Div(lhs.inputs[0], shifted))
    return None


def _cmp_of_div(ranges, op):
    if not isinstance(op, CmpOp) or isinstance(op, Equals):
        return None
    lhs, rhs = op.inputs
    cls = op.__class__
    if isinstance(lhs, Div):
        num, den = lhs.inputs
        return cls(num, den) if _is_one(rhs) else cls(num, den * rhs)
    if isinstance(rhs, Div):
        num, den = rhs.inputs
        return cls(den, num) if _is_one(lhs) else cls(lhs * den, num)
    return None


_RANK_TRANSPARENT = (Sqrt, Log, Scale, Rank)
_POSITIVE_SHIFT = (AddConst, MulConst)
_POSITIVE_UNSHIFT = (SubConst, DivConst)


def _rank_of_increasing(ranges, op):
    if not isinstance(op, Rank):
        return None
    inner = op.inputs[0]
    cls = inner.__class__
    if cls in _RANK_TRANSPARENT:
        return Rank(inner.inputs[0])
    if cls in _POSITIVE_UNSHIFT and not inner.attrs.get("swap") and inner.attrs["value"] > 0:
        return Rank(inner.inputs[0])
    if cls in _POSITIVE_SHIFT and inner.attrs["value"] > 0:
        return Rank(inner.inputs[0])
    if isinstance(inner, Pow) and isinstance(inner.inputs[1], ConstantOp) and inner.inputs[1].attrs["value"] > 0:
        if ranges.get(inner.inputs[0]).start >= 0:
            return Rank(inner.inputs[0])
    return None


def _sign_of_scaled(ranges, op):
    if not isinstance(op, Sign):
        return None
    inner = op.inputs[0]
    if isinstance(inner, Scale):
        return Sign(inner.inputs[0])
    if isinstance(inner, MulConst) and inner.attrs["value"] > 0:
        return Sign(inner.inputs[0])
    if isinstance(inner, DivConst) and inner.attrs["value"] > 0 and not inner.attrs.get("swap"):
        return Sign(inner.inputs[0])
    return None


def _abs_of_nonnegative(ranges, op):
    if isinstance(op, Abs) and ranges.get(op.inputs[0]).start >= 0:
        return op.inputs[0]
    return None


_REWRITES = (_mul_one, _log_difference, _rank_of_increasing, _cmp_of_div, _sign_of_scaled,
             _abs_of_nonnegative)


def _window_sum_source(op: OpBase, enabled: bool):
    if not enabled or not isinstance(op, ReduceAdd) or op.get_parent() is not None:
        return None
    itr = op.inputs[0]
    if not isinstance(itr, IterValue):
        return None
    loop, data = itr.inputs
    if not isinstance(loop, ForeachBackWindow):
        return None
    return data, loop.attrs["window"]


def _rewrite_once(ops: List[OpBase], options: dict):
    replaced: Dict[OpBase, OpBase] = {}
    out: List[OpBase] = []
    ranges = _Ranges()
    changed = False
    for op in ops:
        op.replace_inputs(replaced)
        done = False
        for rule in _REWRITES:
            b = Builder(op.get_parent())
            with b:
                new = rule(ranges, op)
            if new is None:
                continue
            out.extend(b.ops)
            if not b.ops or new != b.ops[-1]:
                out.append(new)
            replaced[op] = new
            changed = done = True
            break
        if done:
            continue
        found = _window_sum_source(op, options.get("opt_reduce", True))
        if found is None:
            out.append(op)
            continue
        data, window = found
        fast = FastWindowedSum(data, window)
        if isinstance(data, WindowedTempOutput):
            assert data.attrs["window"] >= window
            if data.attrs["window"] < window + 1:
                data.attrs["window"] = window + 1
        out.append(fast)
        replaced[op] = fast
        changed = True
    return out if changed else None


@kun_pass
def special_optimize(f, options: dict = {}):
    from ..Stage import Function
    while True:
        newops = _rewrite_once(f.ops, options)
        if newops is None:
            break
        f.set_ops(Function.topo_sort_ops(newops))
