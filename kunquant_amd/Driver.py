"""Compiler driver: option validation, the optimisation pipeline and the
module tables (buffers, stages, dependencies) the device runtime consumes.

Restates KunQuant/Driver.py: KunCompilerConfig 14-23, optimize 25-38,
post_optimize 40-48, option defaults and checks 91-144, buffer naming and
layouts 146-181, per-stage buffer lists 215-245, stage table 246-337.
Instead of C++ source, `compileit` returns a ModuleIR; the HIP emitter
(kunquant_amd.codegen.hip) lowers it to kernels and the oracle consumes its
program text (kunquant_amd.codegen.program)."""
from collections import OrderedDict
from dataclasses import dataclass, field
from typing import Dict, List, Optional, Union

from .Op import CrossSectionalOp, Input, OpBase, Output
from .Stage import Function
from .passes import (decompose, decompose_rank, do_partition, expr_fold, infer_input_window,
                     infer_window, merge_loops, move_dup_rank_output, special_optimize,
                     temp_window_elim)
from .passes import Util as PassUtil

# MI355X backend: blocking_len only defines the STs layout (and the optional
# aligned-shape check); the defaults match the x86 reference so that STs
# arrays produced for the reference are accepted unchanged.
_SUGGESTED_BLOCKING = {"float": 8, "double": 4}
_ALLOWED_BLOCKING = {"float": {8, 16}, "double": {4, 8}}


@dataclass
class KunCompilerConfig:
    partition_factor: int = 3
    dtype: str = "float"
    blocking_len: int = None
    input_layout: str = "STs"
    output_layout: str = "STs"
    allow_unaligned: Union[bool, None] = None
    split_source: int = 0
    options: dict = field(default_factory=dict)


def optimize(f: Function, options: dict) -> Dict[str, int]:
    if PassUtil.debug_mode:
        print("Before optimize: ", f)
    windows = infer_window(f, options)
    special_optimize(f, options)
    decompose(f, options)
    expr_fold(f, options)
    special_optimize(f, options)
    expr_fold(f, options)
    decompose_rank(f, options)
    move_dup_rank_output(f, options)
    temp_window_elim(f, options)
    return windows


def post_optimize(impl: List[Function], options: dict) -> Dict[str, int]:
    windows: Dict[str, int] = {}
    for f in impl:
        temp_window_elim(f, options)
        infer_input_window(f, windows)
        merge_loops(f, options)
    return windows


BUF_INPUT, BUF_OUTPUT, BUF_TEMP = "INPUT", "OUTPUT", "TEMP"
STAGE_TS, STAGE_RANK, STAGE_SCALE, STAGE_GENERIC = "ts", "rank", "scale", "generic"


@dataclass
class BufferIR:
    idx: int
    name: str
    kind: str
    num_users: int = 0
    unreliable: int = 0
    window: int = 1


@dataclass
class StageIR:
    idx: int
    name: str
    kind: str
    func: Function
    in_bufs: List[BufferIR]
    out_bufs: List[BufferIR]
    deps: List[int] = field(default_factory=list)
    dependers: List[int] = field(default_factory=list)


@dataclass
class ModuleIR:
    name: str
    dtype: str
    blocking_len: int
    input_layout: str
    output_layout: str
    aligned: bool
    options: dict
    buffers: List[BufferIR]
    stages: List[StageIR]
    mainf: Function = None

    @property
    def stream_mode(self) -> bool:
        return self.output_layout == "STREAM"

    def buffer(self, name: str) -> BufferIR:
        for b in self.buffers:
            if b.name == name:
                return b
        raise KeyError(name)

    def input_names(self) -> List[str]:
        return [b.name for b in self.buffers if b.kind == BUF_INPUT]

    def output_names(self) -> List[str]:
        return [b.name for b in self.buffers if b.kind == BUF_OUTPUT]


def _first_is_cross_sectional(f: Function) -> Optional[OpBase]:
    for op in f.ops:
        if isinstance(op, (Input, Output)):
            continue
        return op if isinstance(op, CrossSectionalOp) else None
    return None


def _layout_name(name: str, argname: str) -> str:
    if name == "ST8s":
        print(f"The layout name given in {argname} ST8s is depracated. Use STs and blocking_len=8 instead")
        return "STs"
    return name


def compileit(f: Function, module_name: str, partition_factor=3, dtype="float", blocking_len=None,
              input_layout="STs", output_layout="STs", allow_unaligned: Union[bool, None] = None,
              split_source=0, options={}) -> ModuleIR:
    """Optimise and partition `f` and build the module tables. Raises the
    same configuration errors as the reference driver."""
    from .ops.MiscOp import GenericCrossSectionalOp
    options = dict(options)
    input_layout = _layout_name(input_layout, "input_layout")
    output_layout = _layout_name(output_layout, "input_layout")
    if dtype not in ("float", "double"):
        raise RuntimeError("Bad dtype " + dtype)
    if blocking_len is None:
        blocking_len = _SUGGESTED_BLOCKING[dtype]
    if blocking_len not in _ALLOWED_BLOCKING[dtype]:
        raise RuntimeError(f"Blocking length {blocking_len} is not supported for {dtype} on gfx950")
    options["blocking_len"] = blocking_len
    if output_layout not in ("STs", "TS", "STREAM"):
        raise RuntimeError("Bad output_layout name " + output_layout)
    if input_layout not in ("STs", "TS", "STREAM"):
        raise RuntimeError("Bad input_layout name " + input_layout)
    stream_mode = output_layout == "STREAM"
    if stream_mode and input_layout != "STREAM":
        print("Ignoring input_layout because output_layout is stream mode")
        input_layout = "STREAM"
    if stream_mode and options.get("opt_reduce", False):
        raise RuntimeError("Currently opt_reduce in stream mode is not supported.")
    no_warn = False
    if options.get("no_fast_stat", False) == "no_warn":
        options["no_fast_stat"] = False
        no_warn = True
    if stream_mode:
        if "no_fast_stat" not in options:
            options["no_fast_stat"] = True
        elif not options["no_fast_stat"]:
            raise RuntimeError("no_fast_stat=False is not supported in stream mode.")
    else:
        if "no_fast_stat" not in options:
            options["no_fast_stat"] = dtype == "float"
        if not options["no_fast_stat"] and not no_warn:
            print(f"Warning: fast stat optimization is ON for {module_name}. This may result in lower "
                  f"precision and faster execution in windowed statistics functions. You can turn it off "
                  f"by setting options['no_fast_stat'] = True. If you are sure about the precision, you "
                  f"can set options['no_fast_stat'] = 'no_warn' to disable this warning.")
    if stream_mode and allow_unaligned is None:
        allow_unaligned = False
    elif allow_unaligned is None:
        allow_unaligned = True
    if allow_unaligned and stream_mode:
        raise RuntimeError("Currently allow_unaligned in stream mode is not supported.")

    buffers: List[BufferIR] = []
    by_name: Dict[str, BufferIR] = {}

    def buffer_for(name: str, kind: str) -> BufferIR:
        b = by_name.get(name)
        if b is None:
            b = BufferIR(len(buffers), name, kind)
            buffers.append(b)
            by_name[name] = b
        return b

    for op in f.ops:
        if isinstance(op, Input):
            buffer_for(op.attrs["name"], BUF_INPUT)
        elif isinstance(op, Output):
            buffer_for(op.attrs["name"], BUF_OUTPUT)

    unreliable = optimize(f, options)
    mainf, impl = do_partition(f, partition_factor, options)
    windows = post_optimize(impl, options)

    def set_layout(op: OpBase, b: BufferIR):
        if stream_mode:
            op.attrs["layout"] = "STREAM"
        elif op.attrs.get("single_value", False):
            op.attrs["layout"] = "TS"
        elif b.kind == BUF_TEMP:
            op.attrs["layout"] = "TS"
        elif b.kind == BUF_INPUT:
            op.attrs["layout"] = input_layout
        else:
            op.attrs["layout"] = output_layout

    stages: List[StageIR] = []
    by_stage_name: Dict[str, StageIR] = OrderedDict()
    for func in impl:
        ins, outs = [], []
        for op in func.ops:
            if isinstance(op, Input):
                b = buffer_for(op.attrs["name"], BUF_TEMP)
                ins.append(b)
                set_layout(op, b)
            elif isinstance(op, Output):
                b = buffer_for(op.attrs["name"], BUF_TEMP)
                outs.append(b)
                set_layout(op, b)
        for b in ins:
            b.num_users += 1
        cs = _first_is_cross_sectional(func)
        if cs is None:
            kind = STAGE_TS
            for op in func.ops:
                if isinstance(op, CrossSectionalOp):
                    raise RuntimeError(f"Cross sectional op must be alone in its stage: {op}")
        elif isinstance(cs, GenericCrossSectionalOp):
            # its body is compiled as device code (one thread per row)
            kind = STAGE_GENERIC
        else:
            kind = STAGE_RANK if cs.__class__.__name__ == "Rank" else STAGE_SCALE
        st = StageIR(len(stages), func.name, kind, func, ins, outs)
        stages.append(st)
        by_stage_name[func.name] = st
    for p in mainf.ops:
        st = by_stage_name[p.attrs["name"]]
        st.deps = [by_stage_name[i.attrs["name"]].idx for i in p.inputs]
        st.dependers = [by_stage_name[u.attrs["name"]].idx for u in mainf.op_to_id[p].uses]
    for b in buffers:
        b.unreliable = unreliable.get(b.name, 0)
        b.window = windows.get(b.name, 1)
    return ModuleIR(module_name, dtype, blocking_len, input_layout, output_layout, not allow_unaligned,
                    options, buffers, stages, mainf)
