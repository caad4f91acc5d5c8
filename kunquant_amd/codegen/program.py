"""Flat text form of a compiled module ("KUNPROG"): the buffer table, the
stage table and, per time-series stage, its ops in execution order.

The same text drives the CPU oracle (oracle/kun_oracle.c) and is embedded
in every device module for diagnostics. Constants are carried as the
decimal string the reference emits for its C++ literals
(KunQuant/passes/CodegenCpp.py:60-72), so a float module rounds them
decimal->float exactly as the reference compiler does."""
from typing import Dict, List

from ..Driver import BUF_INPUT, BUF_OUTPUT, BUF_TEMP, ModuleIR, STAGE_GENERIC, STAGE_RANK, STAGE_SCALE, STAGE_TS
from ..Op import (ConstantOp, ForeachBackWindow, Input, OpBase, Output, ReductionOp,
                  WindowedTempOutput)

_KIND_CODE = {BUF_INPUT: 0, BUF_OUTPUT: 1, BUF_TEMP: 2}
_STAGE_CODE = {STAGE_TS: 0, STAGE_RANK: 1, STAGE_SCALE: 2}
# generic cross-sectional stages: the ones the oracle knows by name, and 9
# for a user body the oracle cannot run
_GENERIC_CODE = {"DiffWithWeightedSum": 3}


def stage_code(stage) -> int:
    if stage.kind != STAGE_GENERIC:
        return _STAGE_CODE[stage.kind]
    op = next(o for o in stage.func.ops if not isinstance(o, (Input, Output)))
    return _GENERIC_CODE.get(op.__class__.__name__, 9)
_LAYOUT_CODE = {"STs": 0, "TS": 1, "STREAM": 2}


def literal(v) -> str:
    """Decimal text of a constant as the reference prints it."""
    s = str(v)
    if s in ("inf", "-inf"):
        return s
    if "nan" in s.lower():
        return "nan"
    return s


def op_value(op: OpBase) -> str:
    for key in ("value", "q"):
        if key in op.attrs:
            return literal(op.attrs[key])
    return "-"


def op_window(op: OpBase, loop_window: Dict[OpBase, int]) -> int:
    if "window" in op.attrs:
        return int(op.attrs["window"])
    if op.__class__.__name__ == "WindowLoopIndex":
        return int(op.inputs[0].attrs["window"])
    return 0


def stage_ops(module: ModuleIR, stage) -> List[str]:
    f = stage.func
    idx = {op: i for i, op in enumerate(f.ops)}
    lines = []
    for op in f.ops:
        name = op.__class__.__name__
        parent = op.get_parent()
        if isinstance(op, ReductionOp):
            # a reduction lives after its loop; remember which loop feeds it
            parent_code = idx[op.get_loop()]
            name_field = name
            extra_parent = parent_code
        else:
            extra_parent = -1
        buf = -1
        single = 0
        if isinstance(op, (Input, Output)):
            buf = module.buffer(op.attrs["name"]).idx
            single = 1 if op.attrs.get("single_value", False) else 0
        swap = 1 if op.attrs.get("swap", False) else 0
        fields = [
            "O", name, str(idx[parent]) if parent is not None else "-1", str(extra_parent),
            str(len(op.inputs)), *[str(idx[i]) for i in op.inputs], str(buf),
            str(op_window(op, {})), str(swap), str(single), op_value(op),
        ]
        lines.append(" ".join(fields))
    return lines


def program_text(module: ModuleIR) -> str:
    out = ["KUNPROG 1", f"module {module.name}", f"dtype {module.dtype}",
           f"layout {_LAYOUT_CODE[module.input_layout]} {_LAYOUT_CODE[module.output_layout]}",
           f"blocking {module.blocking_len}", f"aligned {1 if module.aligned else 0}",
           f"fastlog {1 if module.options.get('fast_log', True) else 0}",
           f"buffers {len(module.buffers)}"]
    for b in module.buffers:
        out.append(f"B {b.idx} {_KIND_CODE[b.kind]} {b.window} {b.unreliable} {b.name}")
    out.append(f"stages {len(module.stages)}")
    for st in module.stages:
        ops = stage_ops(module, st) if st.kind == STAGE_TS else []
        head = ["S", str(st.idx), str(stage_code(st)), str(len(ops)),
                str(len(st.in_bufs)), *[str(b.idx) for b in st.in_bufs],
                str(len(st.out_bufs)), *[str(b.idx) for b in st.out_bufs],
                str(len(st.deps)), *[str(d) for d in st.deps]]
        out.append(" ".join(head))
        out.extend(ops)
    out.append("end")
    return "\n".join(out) + "\n"
