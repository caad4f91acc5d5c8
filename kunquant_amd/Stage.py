"""`Function`: an ordered, verified, garbage-collected list of ops.

Behavioural restatement of KunQuant/Stage.py:
  * topological sort with loop-body dependencies -> Stage.py:27-59
  * dead-op elimination                          -> Stage.py:61-83
  * set_ops verification                         -> Stage.py:85-108
  * printed form "vN = Cls@{attrs}(vA,vB) in vL"  -> Stage.py:110-125
"""
from collections import OrderedDict
from dataclasses import dataclass
from typing import Dict, List
import typing

from .Op import OpBase, ReductionOp, SinkOpTrait


@dataclass
class OpInfo:
    idx: int
    uses: typing.OrderedDict


class Function:
    def __init__(self, ops: List[OpBase], strict_window: bool = False, name: str = "") -> None:
        self.strict_window = strict_window
        self.name = name
        self.set_ops(ops)

    def get_op_idx(self, op: OpBase) -> int:
        return self.op_to_id[op].idx

    # ------------------------------------------------------------------
    @staticmethod
    def topo_sort_ops(ops: List[OpBase]) -> List[OpBase]:
        """Depth-first topological order. A window loop additionally depends
        on every value its body reads from outside the loop, so that those
        values are computed before the loop starts."""
        present = set(ops)
        loop_extra: Dict[OpBase, Dict[OpBase, None]] = {}
        for op in ops:
            loop = op.get_parent()
            if not loop and isinstance(op, ReductionOp):
                loop = op.get_loop()
            if not loop:
                continue
            extra = loop_extra.setdefault(loop, {})
            for inp in op.inputs:
                if inp != loop and inp.get_parent() != loop:
                    extra[inp] = None

        placed = set()
        order: List[OpBase] = []

        def place(op: OpBase):
            if op not in present:
                raise RuntimeError("dependent Op not in the function: " + str(op))
            if op in placed:
                return
            for dep in loop_extra.get(op, ()):
                place(dep)
            for dep in op.inputs:
                place(dep)
            order.append(op)
            placed.add(op)

        for op in ops:
            place(op)
        return order

    @staticmethod
    def _garbage_collect(ops: List[OpBase], info: Dict[OpBase, OpInfo]) -> List[OpBase]:
        def dead(o):
            return not isinstance(o, SinkOpTrait) and len(info[o].uses) == 0

        worklist = [o for o in ops if dead(o)]
        if not worklist:
            return ops
        while worklist:
            victim = worklist.pop()
            for inp in victim.inputs:
                uses = info[inp].uses
                if victim in uses:
                    del uses[victim]
                    if dead(inp):
                        worklist.append(inp)
        kept = []
        for op in ops:
            if not dead(op):
                kept.append(op)
                info[op].idx = len(kept) - 1
        return kept

    def set_ops(self, ops: List[OpBase]) -> None:
        info: Dict[OpBase, OpInfo] = {}
        for idx, op in enumerate(ops):
            if op in info:
                raise RuntimeError("Duplicated ops: " + str(op))
            for inp in op.inputs:
                if not isinstance(inp, OpBase):
                    raise RuntimeError(f"Bad input: {str(inp)} in {op.__class__.__name__}. op.inputs={op.inputs}")
                if inp not in info:
                    raise RuntimeError(f"Bad op: {str(op)}, because the input has not been executed: {str(inp)}")
                info[inp].uses[op] = 1
            parent = op.get_parent()
            if parent and parent not in info:
                raise RuntimeError(f"Bad op: {str(op)}, because the input has not been executed: {str(parent)}")
            info[op] = OpInfo(idx, OrderedDict())
        self.op_to_id = info
        for op in ops:
            op.verify(self)
        self.ops = Function._garbage_collect(ops, info)
        self.op_to_id = info

    # ------------------------------------------------------------------
    def single_op_str(self, idx: int) -> str:
        op = self.ops[idx]
        args = ",".join(f"v{self.get_op_idx(i)}" for i in op.inputs)
        loop = op.get_parent()
        suffix = f" in v{self.get_op_idx(loop)}" if loop else ""
        return f"v{idx} = {op.__class__.__name__}@{op.attrs_str()}({args}){suffix}"

    def __str__(self) -> str:
        lines = [f"name = {self.name}"] if self.name else []
        lines.extend(self.single_op_str(i) for i in range(len(self.ops)))
        return "\n".join(lines)
