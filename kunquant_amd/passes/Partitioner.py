"""Greedy partitioning of the optimised graph into stages
(KunQuant/passes/Partitioner.py:12-361).

A stage is a fused time loop over a block of stocks; every cross-sectional
op gets a stage of its own (with an Output that directly consumes it).
Selection priority of ready ops, highest first:
  (loop/state/bool-related op, blocks a loop op already in the stage,
   score = 100000 if connected + 50000 per blocked edge + 10 * readiness
           batch - number of transitive users)
with later ready ops winning ties. A stage is closed after more than
`partition_factor` outputs once nothing critical is pending. Values that
cross stages are materialised through Outputs named by their structural
hash. Stage membership does not change numerical results."""
from collections import OrderedDict
from dataclasses import dataclass
from typing import Dict, List, Set, Tuple
import typing

from ..Op import (BoolOpTrait, ConstantOp, CrossSectionalOp, GlobalStatefulProducerTrait,
                  GraphSourceTrait, Input, OpBase, Output, ReductionOp, StateConsumerTrait)
from ..ops.MiscOp import GenericPartition, ReturnFirstValue
from ..Stage import Function
from .Util import debug_mode


@dataclass
class _Info:
    depender: Set[OpBase]
    pending_dep: Dict[OpBase, None]


def _pinned_inside(op: OpBase) -> bool:
    """Ops whose value cannot leave a stage (loop bodies, producer state,
    boolean masks)."""
    return op.get_parent() is not None or isinstance(op, (GlobalStatefulProducerTrait, BoolOpTrait))


def _wants_early(op: OpBase) -> bool:
    if isinstance(op, (ReductionOp, StateConsumerTrait, ReturnFirstValue)) or _pinned_inside(op):
        return True
    return any(isinstance(i, BoolOpTrait) for i in op.inputs)


class _Part:
    def __init__(self):
        self.ops: typing.OrderedDict[OpBase, None] = OrderedDict()
        self.depender: Set[OpBase] = set()
        self.num_outputs = 0
        self.depending = None
        self.stage_op = None
        self.impl_func = None

    def add(self, info, op: OpBase):
        self.ops[op] = None
        if not isinstance(op, GraphSourceTrait):
            if info is not None:
                self.depender.update(info[op].depender)
            if isinstance(op, Output):
                self.num_outputs += 1

    def reaches_outside(self, op: OpBase) -> bool:
        return op not in self.ops and op in self.depender

    def edge_ops(self, f: Function) -> Dict[OpBase, bool]:
        edges: Dict[OpBase, bool] = {}
        for op in self.ops:
            if isinstance(op, GraphSourceTrait):
                continue
            pinned = _pinned_inside(op)
            for user in f.op_to_id[op].uses:
                if user not in self.ops:
                    edges[user] = pinned or edges.get(user, False)
        return edges


def _collect(f: Function) -> Dict[OpBase, _Info]:
    info: Dict[OpBase, _Info] = {}
    for op in reversed(f.ops):
        dep = set()
        for user in f.op_to_id[op].uses:
            dep.add(user)
            dep.update(info[user].depender)
        info[op] = _Info(dep, dict((i, None) for i in op.inputs))
    return info


def _check_add(op: OpBase, part: _Part) -> Tuple[bool, bool]:
    connected = bad = False
    for inp in op.inputs:
        if inp in part.ops:
            connected = True
        elif part.reaches_outside(inp):
            bad = True
    return bad, connected


def _pick(ready, info, part: _Part, f: Function, edges):
    best = (-1, -1, -1)
    best_op = None
    if edges is None:
        edges = part.edge_ops(f)
    for op, batch in ready:
        if isinstance(op, CrossSectionalOp):
            return op
        bad, connected = _check_add(op, part)
        if bad:
            continue
        deps = info[op].depender
        critical = 0
        score = 0
        for edge_op, pinned in edges.items():
            if edge_op in deps:
                if pinned:
                    critical = 2 if isinstance(op, Output) else 3
                    break
                score += 50000
        if connected:
            score += 100000
        score += batch * 10 - len(deps)
        key = (1 if _wants_early(op) else 0, critical, score)
        if key >= best:
            best, best_op = key, op
    return best_op


def _partition(f: Function, threshold=3) -> List[_Part]:
    info = _collect(f)
    parts: List[_Part] = []
    ready = [(op, 0) for op in f.ops if isinstance(op, GraphSourceTrait)]
    unvisited = dict((op, None) for op in f.ops)
    batch = 0

    def retire(op: OpBase):
        nonlocal batch
        batch += 1
        for user in f.op_to_id[op].uses:
            pending = info[user].pending_dep
            del pending[op]
            if not pending:
                ready.append((user, batch))
        del unvisited[op]
        for i, (o, _) in enumerate(ready):
            if o == op:
                del ready[i]
                break
        else:
            raise RuntimeError("del_from_ready_op: OP not found")

    while ready:
        part = _Part()
        chosen = _pick(ready, info, part, f, None)
        while chosen:
            retire(chosen)
            if isinstance(chosen, GraphSourceTrait):
                pass
            elif isinstance(chosen, CrossSectionalOp):
                solo = _Part()
                solo.add(info, chosen)
                for user in f.op_to_id[chosen].uses:
                    if isinstance(user, Output):
                        retire(user)
                        solo.add(info, user)
                parts.append(solo)
            else:
                for inp in chosen.inputs:
                    if isinstance(inp, GraphSourceTrait):
                        part.add(info, inp)
                part.add(info, chosen)
            edges = part.edge_ops(f)
            if part.num_outputs > threshold and not any(edges.values()):
                direct = None
                for cand, _ in ready:
                    if isinstance(cand, Output) and not _check_add(cand, part)[0]:
                        direct = cand
                        break
                if direct is not None:
                    chosen = direct
                    continue
                break
            chosen = _pick(ready, info, part, f, edges)
        if part.ops:
            parts.append(part)
    if unvisited:
        raise RuntimeError("Some ops are not visited")
    return parts


def _output_of(op: OpBase, info):
    if isinstance(op, Output):
        return op
    for use in info.uses:
        if isinstance(use, Output):
            return use
    return None


def _print_partition_info(stages, impl):
    if debug_mode < 2:
        return
    by_name = dict((fn.name, fn) for fn in impl)
    for s in stages:
        print(f"===========================\npartition: {s.attrs['name']}\n"
              f"inputs: {[i.attrs['name'] for i in s.inputs]}\nImpl:\n{by_name[s.attrs['name']]}")


def _transform_partitions(parts: List[_Part], f: Function):
    names: Dict[str, object] = {}
    for op in f.ops:
        if isinstance(op, (Input, Output)):
            name = op.attrs["name"]
            if name == "" or name in names:
                raise RuntimeError(f"Duplicated or bad name of op: {id(op)} {op}\n"
                                   f"old op is: {id(names.get(name))} {names.get(name)}")
            names[name] = op

    def fresh_name(base: str) -> str:
        cur, idx = base, 0
        while cur in names:
            cur = f"{base}_{idx}"
            idx += 1
        names[cur] = None
        return cur

    owner: Dict[OpBase, _Part] = {}
    for p in parts:
        for op in p.ops:
            if not isinstance(op, GraphSourceTrait):
                assert op not in owner
                owner[op] = p
    hash_cache: Dict[OpBase, int] = {}
    for p in parts:
        local_inputs: Dict[str, Input] = {}
        depending: typing.OrderedDict = OrderedDict()
        for op in list(p.ops):
            for idx, inp in enumerate(op.inputs):
                if inp in p.ops:
                    continue
                if inp.get_parent():
                    raise RuntimeError("Bad cross partition op: " + str(inp) + "\ncur op=" + str(op))
                inp_info = f.op_to_id[inp]
                if isinstance(inp, ConstantOp):
                    inp_info.uses.pop(op, None)
                    clone = ConstantOp(inp.attrs["value"])
                    p.add(None, clone)
                    op.inputs[idx] = clone
                    continue
                if isinstance(inp, Input):
                    buf_name = inp.attrs["name"]
                else:
                    out = _output_of(inp, inp_info)
                    if not out:
                        buf_name = fresh_name(inp.hash_hex(hash_cache))
                        out = Output(inp, buf_name)
                        inp_info.uses[out] = 1
                        src_part = owner[inp]
                        src_part.add(None, out)
                        owner[out] = src_part
                    else:
                        buf_name = out.attrs["name"]
                        src_part = owner[out]
                    if src_part != p:
                        depending[src_part] = None
                inp_info.uses.pop(op, None)
                local = local_inputs.get(buf_name)
                if not local:
                    local = Input(buf_name)
                    p.add(None, local)
                    local_inputs[buf_name] = local
                op.inputs[idx] = local
        p.depending = depending
        p.stage_op = GenericPartition([], None)

    impls, stages = [], []
    for p in parts:
        ops = Function.topo_sort_ops(list(p.ops))
        name = "_".join(op.attrs["name"] for op in ops if isinstance(op, Output))
        p.impl_func = Function(ops, True, name)
        p.stage_op.attrs["name"] = name
        stages.append(p.stage_op)
        impls.append(p.impl_func)
        for dep in p.depending:
            p.stage_op.inputs.append(dep.stage_op)
    _print_partition_info(stages, impls)
    return Function(Function.topo_sort_ops(stages)), impls


def do_partition(f: Function, factor: int, options: dict = {}):
    return _transform_partitions(_partition(f, factor), f)
