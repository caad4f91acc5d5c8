"""Decomposition of compositive ops, window-source insertion and
materialisation of cross-sectional inputs.

Restates KunQuant/passes/Decompose.py: decompose_impl 7-36,
decompose_rank_impl 38-75, move_dup_rank_output_impl 90-126."""
from typing import Dict, List

from ..Op import (CompositiveOp, CrossSectionalOp, Input, OpBase, Output, WindowedDataSourceOp,
                  WindowedTempOutput, WindowedTrait)
from .Util import kun_pass


def _decompose_list(ops: List[OpBase], options) -> List[OpBase]:
    replaced: Dict[OpBase, OpBase] = {}
    out: List[OpBase] = []
    changed = False
    for op in ops:
        op.replace_inputs(replaced)
        if isinstance(op, CompositiveOp):
            changed = True
            parts = op.decompose(options)
            deeper = _decompose_list(parts, options)
            if deeper is not None:
                parts = deeper
            out.extend(parts)
            replaced[op] = parts[-1]
        elif isinstance(op, WindowedTrait):
            window = op.required_input_window()
            for idx, inp in enumerate(op.inputs):
                if not isinstance(inp, WindowedDataSourceOp):
                    changed = True
                    tmp = WindowedTempOutput(inp, window)
                    op.inputs[idx] = tmp
                    out.append(tmp)
            out.append(op)
        else:
            out.append(op)
    return out if changed else None


def _materialize_cs_inputs(ops: List[OpBase], info) -> List[OpBase]:
    replaced: Dict[OpBase, OpBase] = {}
    out: List[OpBase] = []
    changed = False
    made: Dict[OpBase, OpBase] = {}
    seen = set()
    hash_cache: Dict[OpBase, int] = {}
    for op in ops:
        op.replace_inputs(replaced)
        if op in seen:
            continue
        if isinstance(op, CrossSectionalOp):
            for idx, inp in enumerate(op.inputs):
                if isinstance(inp, (Input, Output)):
                    continue
                changed = True
                target = made.get(inp)
                if target is None:
                    for user in info[inp].uses:
                        if isinstance(user, Output):
                            target = user
                            if user not in seen:
                                seen.add(user)
                                out.append(user)
                            break
                    if target is None:
                        target = Output(inp, inp.hash_hex(hash_cache))
                        out.append(target)
                    made[inp] = target
                op.inputs[idx] = target
        out.append(op)
        seen.add(op)
    return out if changed else None


@kun_pass
def decompose(f, options: dict = {}):
    newops = _decompose_list(f.ops, options)
    f.strict_window = True
    if newops is not None:
        f.set_ops(newops)


@kun_pass
def decompose_rank(f, options: dict = {}):
    newops = _materialize_cs_inputs(f.ops, f.op_to_id)
    if newops is not None:
        f.set_ops(newops)


def _share_cs_outputs(ops: List[OpBase]):
    """A second Output of the same cross-sectional value reads the first
    Output instead of the value itself (in place; never reports a change)."""
    first: Dict[OpBase, OpBase] = {}
    edited = False
    for op in ops:
        if isinstance(op, Output):
            src = op.inputs[0]
            if isinstance(src, CrossSectionalOp):
                if src in first:
                    op.inputs[0] = first[src]
                    edited = True
                else:
                    first[src] = op
    return edited


@kun_pass
def move_dup_rank_output(f, options: dict = {}):
    # The op list is unchanged; the use lists are rebuilt so that later
    # passes see the re-pointed Output (the reference leaves them stale).
    if _share_cs_outputs(f.ops):
        f.set_ops(list(f.ops))
