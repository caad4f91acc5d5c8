"""Sharing of window loops over one buffer (KunQuant/passes/MergeLoops.py).

Loops of different windows over the same buffer, each feeding exactly one
single-input reduction, are chained longest-first: a loop iterates from its
window down to the next loop's window (segment_end) and the next loop
re-runs the previous bodies (copy_prev_body). Every reduction still sees
its own window oldest-to-newest. Loops with equal windows share one body."""
from typing import Dict, List, Optional

from ..Op import ForeachBackWindow, OpBase, ReductionOp
from .Util import kun_pass


def _only_user(f, op: OpBase) -> Optional[OpBase]:
    uses = f.op_to_id[op].uses
    if len(uses) != 1:
        return None
    return next(iter(uses))


def _loop_groups(f, op: OpBase):
    by_window: Dict[int, List[OpBase]] = {}
    for loop in f.op_to_id[op].uses:
        if len(loop.inputs) != 1 or not isinstance(loop, ForeachBackWindow):
            continue
        itr = _only_user(f, loop)
        if not itr:
            continue
        red = _only_user(f, itr)
        if not red or not isinstance(red, ReductionOp) or len(red.inputs) != 1:
            continue
        w = loop.attrs["window"]
        if w in by_window:
            by_window[w].append(red)
        else:
            by_window[w] = [loop, itr, red]
    if len(by_window) <= 1:
        return None
    return sorted(by_window.values(), key=lambda g: g[0].attrs["window"], reverse=True)


def _merge(f, ops: List[OpBase]):
    consumed = set()
    out: List[OpBase] = []
    changed = False
    for op in ops:
        if op in consumed:
            continue
        groups = _loop_groups(f, op)
        out.append(op)
        if not groups:
            continue
        changed = True
        for gi, group in enumerate(groups):
            loop, itr = group[0], group[1]
            if gi != 0:
                loop.attrs["copy_prev_body"] = True
            if gi != len(groups) - 1:
                loop.attrs["segment_end"] = groups[gi + 1][0].attrs["window"]
            for red in group[3:]:
                red.inputs[0] = itr
            for o in group:
                out.append(o)
                consumed.add(o)
    return out if changed else None


@kun_pass
def merge_loops(f, options: dict = {}):
    newops = _merge(f, f.ops)
    if newops is not None:
        f.set_ops(newops)
