"""Window-buffer sharing (KunQuant/passes/TempWindowElim.py:12-62).

A WindowedTempOutput is dropped when its source is a graph Input, already
has an Output (whose buffer holds the full history), or has another
WindowedTempOutput with at least the same window."""
from typing import Dict, List, Optional

from ..Op import Input, OpBase, Output, WindowedTempOutput, traverse_replace_map
from .Util import kun_pass


def _replacement(op: OpBase, f, replaced: Dict[OpBase, OpBase]) -> Optional[OpBase]:
    if not isinstance(op, WindowedTempOutput):
        return None
    src = op.inputs[0]
    if isinstance(src, Input):
        return src
    need = op.attrs["window"]
    best, best_w = None, 0
    for user in f.op_to_id[src].uses:
        if user == op:
            continue
        if isinstance(user, Output):
            return traverse_replace_map(user, replaced)
        if isinstance(user, WindowedTempOutput):
            w = user.attrs["window"]
            # ties between equal windows go to the later user
            if w >= need and w >= best_w and (w > best_w or best is not None):
                best, best_w = user, w
    if best is not None:
        return traverse_replace_map(best, replaced)
    return None


def _elim(ops: List[OpBase], f):
    replaced: Dict[OpBase, OpBase] = {}
    out: List[OpBase] = []
    changed = False
    for op in ops:
        if op in replaced:
            continue
        op.replace_inputs(replaced)
        repl = _replacement(op, f, replaced)
        if repl is None:
            out.append(op)
        else:
            changed = True
            replaced[op] = repl
    return out if changed else None


@kun_pass
def temp_window_elim(f, options: dict = {}):
    from ..Stage import Function
    newops = _elim(f.ops, f)
    if newops is not None:
        f.set_ops(Function.topo_sort_ops(newops))
