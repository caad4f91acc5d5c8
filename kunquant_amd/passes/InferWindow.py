"""Warm-up length ("unreliable count") and per-buffer history window
inference (KunQuant/passes/InferWindow.py:6-33)."""
from typing import Dict

from ..Op import Input, OpBase, Output, WindowedTrait


def _lookback(op: OpBase, memo: Dict[OpBase, int]) -> int:
    if op in memo:
        return memo[op]
    v = max((_lookback(i, memo) for i in op.inputs), default=0)
    if isinstance(op, WindowedTrait):
        v += op.attrs["window"]
    memo[op] = v
    return v


def infer_window(f, options: dict = {}) -> Dict[str, int]:
    """Sum of windows along the longest path to every Output."""
    memo: Dict[OpBase, int] = {}
    return {op.attrs["name"]: _lookback(op, memo) for op in f.ops if isinstance(op, Output)}


def infer_input_window(f, result: Dict[str, int]) -> None:
    """Rows of history each named buffer must keep (at least 1)."""
    for op in f.ops:
        if isinstance(op, (Input, Output)):
            name = op.attrs["name"]
            result[name] = result.get(name, 1)
            for use in f.op_to_id[op].uses:
                if isinstance(use, WindowedTrait):
                    result[name] = max(use.required_input_window(), result.get(name, 0))
