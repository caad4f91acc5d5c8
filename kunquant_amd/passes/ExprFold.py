"""Common-subexpression elimination (KunQuant/passes/ExprFold.py:10-57).

Two ops are merged when they live in the same loop, have the same class
and attributes, and either share the same input objects or have identical
structural token streams."""
from typing import Dict, List

from ..Op import ForeachBackWindow, OpBase, Output
from .Util import kun_pass


class _Key:
    __slots__ = ("op", "cache", "loop")

    def __init__(self, op: OpBase, cache: Dict[OpBase, int], loop):
        self.op, self.cache, self.loop = op, cache, loop

    def __hash__(self):
        return self.op.fast_hash(self.cache)

    def __eq__(self, other: "_Key") -> bool:
        a, b = self.op, other.op
        if self.loop != other.loop or a.__class__ != b.__class__ or a.attrs != b.attrs:
            return False
        if a.inputs == b.inputs:
            return True
        for x, y in zip(a.items(), b.items()):
            if x != y:
                return False
        return True


def _fold(ops: List[OpBase]):
    replaced: Dict[OpBase, OpBase] = {}
    out: List[OpBase] = []
    seen: Dict[_Key, OpBase] = {}
    cache: Dict[OpBase, int] = {}
    changed = False
    for op in ops:
        op.replace_inputs(replaced)
        if isinstance(op, (ForeachBackWindow, Output)):
            out.append(op)
            continue
        key = _Key(op, cache, op.get_parent())
        prior = seen.get(key)
        if prior is None:
            seen[key] = op
            out.append(op)
        else:
            replaced[op] = prior
            changed = True
    return out if changed else None


@kun_pass
def expr_fold(f, options: dict = {}):
    newops = _fold(f.ops)
    if newops is not None:
        f.set_ops(newops)
