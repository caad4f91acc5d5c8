"""HIP emitter: lowers the stages of a ModuleIR to gfx950 kernels built from
the device templates in csrc/kun_device.hpp.

Batch time-series stages follow the kernel plan of codegen/plan.py: each
level becomes one kernel over a flat grid of workgroups; the runtime's entry
table maps a workgroup to an independent component of the level, a chunk of
rows (PAR levels; one chunk for SEQ levels) and 64 stocks (one lane per
stock, TS rows coalesced).
  * every value read through a window keeps a per-lane history ring laid
    out [slot][lane] (bank-conflict free), slot count a power of two,
    NaN-initialised -- the reference's "no history before start" rule
    (Ops.hpp:87-92, 182-186). Rings of values computed in the kernel live
    in LDS (HBM scratch past the LDS budget); values the kernel reads from
    HBM (stage inputs, values materialised by a lower level) are ringed in
    LDS while the budget allows and otherwise re-read from HBM;
  * a PAR chunk starts `halo` rows before its first stored row so that every
    window it reads is rebuilt; rings of HBM values are prefilled instead;
  * HBM values are prefetched one or more rows ahead in registers;
  * window loops run oldest -> newest like CodegenCpp.py:242-270.
Stream stage kernel (STREAM modules): the same op lowering for one row;
rings and carried state live in HBM as [slot][S] arrays owned by the stream
context (the reference's StreamBuffer / StateBuffer, Runtime.cpp:384-546).
Cross-sectional stages (Rank / Scale) are runtime library kernels."""
import functools
import math
import os
import re
from typing import Dict, List, Optional, Tuple

import numpy as np

from ..Driver import BUF_INPUT, BUF_TEMP, ModuleIR, STAGE_GENERIC, STAGE_TS
from ..Op import ConstantOp, ForeachBackWindow, Input, OpBase, Output, ReductionOp
from .plan import PAR, SEQ, Component, StagePlan, plan_stage
from .program import literal, program_text
from .wire import wire_table

TS_BLOCK = 64
MAX_STAGE_BUFS = 256
LDS_LIMIT = 160 * 1024
# LDS per 64-stock workgroup before history rings spill to HBM scratch, in
# ring slots (one row of 64 stocks each): 80 slots are 20 KiB in fp32 and
# 40 KiB in fp64, which measured fastest for both on the 100-factor graph
# (more LDS per workgroup costs more in occupancy than the HBM re-reads it
# saves). KUN_LDS_BUDGET gives bytes instead.
LDS_SLOTS = 80
LDS_BUDGET = int(os.environ["KUN_LDS_BUDGET"]) if os.environ.get("KUN_LDS_BUDGET") else None


# Sequential components run one workgroup per 64 stocks, far fewer than the
# GPU holds, so occupancy costs them nothing: their rings get up to 80 KiB
# (two workgroups per CU), and a window loop over a computed value reads
# LDS instead of HBM scratch on every iteration of every row
SEQ_LDS_BUDGET = 80 * 1024


def lds_budget_for(module: ModuleIR, override: Optional[int] = None, seq: bool = False) -> int:
    """LDS bytes per workgroup for `module`'s history rings (`seq`: of a
    sequential component)."""
    if override is not None:
        return override
    if LDS_BUDGET is not None:
        return LDS_BUDGET
    if seq:
        return SEQ_LDS_BUDGET
    return LDS_SLOTS * TS_BLOCK * (8 if module.dtype == "double" else 4)
# registers of prefetched loads per component, and the deepest prefetch
# (rows) of a small sequential component
PREFETCH_REGS = int(os.environ.get("KUN_PREFETCH_REGS", 64))
SEQ_PREFETCH_MAX = int(os.environ.get("KUN_SEQ_PREFETCH_MAX", 16))
# rows unrolled (and prefetched) in sequential components of more than 60
# ops without window loops (1 before: the fp64 100-factor graph's
# accumulator webs then waited for every row's loads; 2 / 4 rows measured
# 77.7 / 77.8 against 78.9 ms, Alpha158 fp64 17.9 / 17.7 against 18.3 ms)
SEQ_HEAVY_UNROLL = 4
# rows unrolled in sequential components with window loops (2 / 4 rows
# measured within 2% of 1 on the fp64 graphs, either way)
SEQ_LOOP_UNROLL = 1
# unroll factor of window loops (10: the common 10/20/30/60-row windows
# split into whole groups; measured best of 6/8/10/12 on the 100-factor
# graph)
LOOP_UNROLL = int(os.environ.get("KUN_LOOP_UNROLL", 10))
SORTED_REGS = 256  # register budget (values) for the sorted windows of one sequential component
# row-chunked components: rows unrolled per group by estimated cost per row
# (first threshold met wins; lighter components unroll 8 rows)
PAR_UNROLL_BY_COST = ((1024, 1), (256, 2), (64, 4))
# scheduling barrier between the unrolled rows of a non-interleaved loop
# (mask of the instruction kinds allowed across it), or None. 0x9C lets
# scalar, memory and LDS instructions cross but no vector ALU: without it
# the scheduler lifts every row's NaN select of its lag value to the top of
# the 16-row iteration, where it waits for loads issued a row earlier
ROW_BARRIER = 0x9C
# row-chunked components with window loops: rows per interleaved group
# (2: at two waves per SIMD, wider groups cost more in registers than they
# gain in overlapped chains; measured against 1, 3 and 4)
INTERLEAVE_MIN, INTERLEAVE_MAX = 2, 2

_BIN = {"Add": "+", "Sub": "-", "Mul": "*", "Div": "/"}
_CMP = {"GreaterThan": ">", "GreaterEqual": ">=", "LessThan": "<", "LessEqual": "<=", "Equals": "=="}
_CONST_BIN = {"AddConst": "+", "SubConst": "-", "MulConst": "*", "DivConst": "/",
              "GreaterThanConst": ">", "LessThanConst": "<"}
_UNARY = {"Sqrt": "vsqrt", "Abs": "vabs", "Sign": "vsign", "Not": "bnot", "Exp": "exp_fast"}
_BINFN = {"Max": "vmax", "Min": "vmin", "Or": "bor", "And": "band"}
_LAG_OPS = ("BackRef", "FastWindowedSum", "WindowedLinearRegression")
# rough VALU instructions per evaluation, for balancing work across workgroups
_OP_COST = {"Div": 10, "DivConst": 10, "Sqrt": 8, "Log": 24, "Exp": 24, "IterValue": 2, "Select": 2,
            "ReduceAdd": 4, "ReduceMin": 4, "ReduceMax": 4, "ReduceArgMax": 7, "ReduceArgMin": 7,
            "ReduceRank": 6, "ReduceDecayLinear": 2, "FastWindowedSum": 16, "ExpMovingAvg": 8,
            "WindowedLinearRegression": 30, "SkipListState": 60, "ForeachBackWindow": 2}
MAX_COMPS = 64
# upper LDS bytes of the launch classes a level's components are split into
# (16, 8, 4, 2 and 1 workgroups of 64 lanes per CU; the 80 KiB class is
# where sequential components with large rings go)
LDS_CLASSES = (0, 10 * 1024, 20 * 1024, 40 * 1024, 80 * 1024, LDS_LIMIT)


def c_literal(v, T: str) -> str:
    """C++ literal for a constant, rounded from the same decimal text as
    the reference's generated code (CodegenCpp.py:60-72)."""
    s = literal(v)
    if s == "inf":
        return f"kun::inf<{T}>()"
    if s == "-inf":
        return f"(-kun::inf<{T}>())"
    if s == "nan":
        return f"kun::qnan<{T}>()"
    if "." not in s and "e" not in s:
        s += "."
    return s + ("f" if T == "float" else "")


def _pow2(n: int) -> int:
    r = 1
    while r < n:
        r <<= 1
    return r


@functools.lru_cache(maxsize=None)
def div_product_exact(c: float) -> bool:
    """True when (float)((double)x * RN(1/c)) is the IEEE float quotient x / c
    for every float x (c a finite nonzero float). The product is within
    2^-52 of x / c, which is never that near a rounding midpoint unless it
    is one exactly (see kun::div_rc); exact midpoints exist only between
    subnormal floats, for c = C * 2^k with C odd and k >= 1: x = o * C *
    2^(k-150) for odd o with o * C < 2^24. Each such x is tried."""
    n, d = float(c).as_integer_ratio()
    n = abs(n)
    if d != 1 or n & 1:
        return True            # k < 1: no float x is a midpoint times c
    k = (n & -n).bit_length() - 1
    C = n >> k
    if C == 1:
        return True            # power of two: the product is exact
    rc = 1.0 / float(c)
    o = np.arange(1, (1 << 24) // C + 1, 2, dtype=np.int64)
    o = o[o * C < (1 << 24)]
    x = (o * C).astype(np.float64) * 2.0 ** (k - 150)     # exact, and floats
    with np.errstate(all="ignore"):
        want = x.astype(np.float32) / np.float32(c)
        got = (x * rc).astype(np.float32)
    return bool(np.array_equal(want.view(np.uint32), got.view(np.uint32)))


def ident(name: str) -> str:
    return "".join(c if c.isalnum() else "_" for c in name)


def _window_reads(op) -> Optional[Tuple[OpBase, int]]:
    """(source, largest offset) of an op reading a value through a window."""
    name = op.__class__.__name__
    if name == "IterValue":
        return op.inputs[1], op.inputs[0].attrs["window"] - 1
    if name in _LAG_OPS:
        return op.inputs[0], op.attrs["window"]
    return None


def _rank_copy_source(rank) -> Optional[OpBase]:
    """The ReduceRank a rank's current value is a copy of, or None. The
    reference builds a rank's state as `ReduceRank<T, N> v{cur}`
    (CodegenCpp.py:246-270); when `cur` is itself a ReduceRank object --
    directly or through Output / WindowedTempOutput / ReturnFirstValue, which
    the reference emits as `auto v = w` copies -- C++ selects the copy
    constructor, so the new rank compares against the other rank's value and
    continues its less / equal counts (Ops.hpp:598-616)."""
    src = rank.inputs[1]
    while src.__class__.__name__ in ("Output", "WindowedTempOutput", "ReturnFirstValue"):
        src = src.inputs[0]
    return src if src.__class__.__name__ == "ReduceRank" else None


_DEF = re.compile(r"\bconst\s+\w+\s+(\w+)\s*=")
_ITER = re.compile(r"\biter\b")


def loop_code(w: int, body: List[str], unroll: int) -> List[str]:
    """A window loop over iter = w-1 .. 0 (oldest value first), unrolled by
    hand in groups of `unroll` iterations (0: the whole window): the group's
    window reads -- statements that depend on nothing but `iter` -- are
    issued first, then the iterations run in order. Every statement of the
    loop body keeps its order relative to the others of its iteration, so
    the arithmetic is that of the plain loop; only the loads move earlier,
    so that their latency overlaps the dependent chains."""
    U = w if unroll <= 0 else min(unroll, w)
    defined = set()
    for x in body:
        defined.update(_DEF.findall(x))
    hoist = [i for i, x in enumerate(body)
             if x.lstrip().startswith("const ") and _ITER.search(x) and not any(
                 re.search(rf"\b{n}\b", x.split("=", 1)[1]) for n in defined)]

    def copy(u: int) -> List[str]:
        out = []
        for x in body:
            x = _ITER.sub(f"(it0 - {u})", x)
            for n in defined:
                x = re.sub(rf"\b{n}\b", f"{n}u{u}", x)
            out.append(x)
        return out

    def group(n: int) -> List[str]:
        copies = [copy(u) for u in range(n)]
        G = []
        for c in copies:
            G.extend(c[i] for i in hoist)
        for c in copies:
            G.extend(c[i] for i in range(len(c)) if i not in hoist)
        return G

    L = ["{", "int it0 = " + str(w - 1) + ";", f"for (; it0 >= {U - 1}; it0 -= {U}) {{"]
    L.extend(group(U))
    L.append("}")
    if w % U:
        # the last w % U iterations as a plain loop (unrolling it like the
        # groups measured no faster in batch kernels and slower in stream
        # tick kernels, whose window reads come from HBM)
        L.append("for (int iter = it0; iter >= 0; --iter) {")
        L.extend(body)
        L.append("}")
    L.append("}")
    return L


def window_loop(w: int, body: List[str], fast: Optional[List[str]], unroll: int, row: str = "i") -> List[str]:
    """loop_code of a window loop; with an in-run body, that copy runs when
    every row the loop reads lies in the run (`row` >= w - 1, `row` being
    the lowest row of the loop) and the buffers allow kun::wld."""
    if fast is None:
        return loop_code(w, body, unroll)
    return ([f"if (wok && {row} >= {w - 1}) {{"] + loop_code(w, fast, unroll) + ["} else {"]
            + loop_code(w, body, unroll) + ["}"])


def _flatten(segs, unroll: int) -> List[str]:
    L = []
    for sg in segs:
        if sg[0] == "lines":
            L.extend(sg[1])
        else:
            L.extend(sg[1] + window_loop(sg[2], sg[3], sg[5], unroll) + sg[4])
    return L


_RENAME = [(re.compile(r"\br(\d+)_"), r"r\1q{q}_"), (re.compile(r"\blg(\d+)_(\d+)\b"), r"lg\1_\2q{q}"),
           (re.compile(r"\blr(\d+)_(\d+)\b"), r"lr\1_\2q{q}"),
           (re.compile(r"\brp(\d+)\b"), r"rp\1q{q}"), (re.compile(r"\bi\b"), "iq{q}")]
_VNAME = re.compile(r"\bv(\d+)\b")


def _rename_row(lines: List[str], q: int, keep: set) -> List[str]:
    """Names of one row's statements made private to row copy q of an
    interleaved group: values v<k> (but those in `keep`, declared outside
    the row), reduction state r<k>_*, lag values, ring positions and the row
    index i."""
    out = []
    for x in lines:
        x = _VNAME.sub(lambda m: m.group(0) if int(m.group(1)) in keep else f"v{m.group(1)}_{q}", x)
        for pat, rep in _RENAME:
            x = pat.sub(rep.replace("{q}", str(q)), x)
        out.append(x)
    return out


class _Lowering:
    """Per-row lowering of stage ops to HIP statements, shared by the batch
    component kernels and the stream tick kernels. Subclasses provide the
    storage hooks: win(), input_expr(), output_lines(), after_value(),
    lane_expr() and tmem()."""

    loop_unroll = LOOP_UNROLL  # window loops: #pragma unroll factor (0: full)

    time_idx = "i"

    def _setup(self, module: ModuleIR, stage):
        self.m = module
        self.st = stage
        self.T = "float" if module.dtype == "float" else "double"
        self.esz = 4 if self.T == "float" else 8
        # option fp64_finalize (fp32 modules): window-loop sums and linear
        # regression sums accumulate in double, rounded to float once at the
        # end -- the steps where sum-of-squares cancellation bites
        self.fin = self.T == "float" and bool(module.options.get("fp64_finalize", False))
        self.AT = "double" if self.fin else self.T
        self.idx: Dict[OpBase, int] = {op: i for i, op in enumerate(stage.func.ops)}
        self.slot: Dict[int, int] = {}
        for b in list(stage.in_bufs) + list(stage.out_bufs):
            if b.idx not in self.slot:
                self.slot[b.idx] = len(self.slot)
        if len(self.slot) > MAX_STAGE_BUFS:
            raise RuntimeError(f"stage {stage.name} uses {len(self.slot)} buffers (limit {MAX_STAGE_BUFS})")
        # the runtime passes inputs then outputs, so materialised values start at len(slot)
        assert len(self.slot) == len(stage.in_bufs) + len(stage.out_bufs), stage.name

    def v(self, op) -> str:
        return f"v{self.idx[op]}"

    def buf_layout(self, op) -> str:
        b = self.m.buffer(op.attrs["name"])
        if b.kind == BUF_TEMP:
            return "1"
        lay = self.m.input_layout if b.kind == BUF_INPUT else self.m.output_layout
        return "1" if lay == "TS" else "0"

    def ring_read(self, k: int, off) -> str:
        slots, base = self.ring[k]
        return self.tmem(f"{base} + (((int)i - ({off})) & {slots - 1})", base)

    def ring_push(self, k: int) -> List[str]:
        if k not in self.ring:
            return []
        slots, base = self.ring[k]
        return [f"{self.tmem(f'{base} + ((int)i & {slots - 1})', base)} = v{k};"]

    def after_value(self, k: int) -> List[str]:
        return self.ring_push(k)

    # -- ops -------------------------------------------------------------
    def op_lines(self, op) -> List[str]:
        T = self.T
        k = self.idx[op]
        name = op.__class__.__name__
        ins = [self.v(i) for i in op.inputs]
        out = f"v{k}"
        L: List[str] = []
        time_idx = self.time_idx
        if isinstance(op, Input):
            L.append(f"const {T} {out} = {self.input_expr(op)};")
        elif isinstance(op, Output):
            L.extend(self.output_lines(op, ins[0]))
            L.append(f"const {T} {out} = {ins[0]};")
        elif name in ("WindowedTempOutput", "ReturnFirstValue"):
            L.append(f"const {T} {out} = {ins[0]};")
        elif isinstance(op, ConstantOp):
            L.append(f"const {T} {out} = {c_literal(op.attrs['value'], T)};")
        elif name in _CONST_BIN:
            c = c_literal(op.attrs["value"], T)
            x, y = (c, ins[0]) if op.attrs.get("swap", False) else (ins[0], c)
            expr = f"({x} {_CONST_BIN[name]} {y})"
            cf = float(np.float32(op.attrs["value"]))
            if name == "DivConst" and T == "float" and not op.attrs.get("swap", False) \
                    and math.isfinite(cf) and cf != 0.0:
                # float division by a constant as a double product, exact:
                # bare where no midpoint quotient exists or all round right
                # (div_product_exact), else subnormal results divide
                fn = "kun::div_const" if div_product_exact(cf) else "kun::div_rc"
                expr = f"{fn}({ins[0]}, {float.hex(1.0 / cf)}, {c})"
            if name.endswith("ThanConst"):
                expr = f"kun::mask<{T}>{expr}"
            L.append(f"const {T} {out} = {expr};")
        elif name in _BIN:
            L.append(f"const {T} {out} = {ins[0]} {_BIN[name]} {ins[1]};")
        elif name in _CMP:
            L.append(f"const {T} {out} = kun::mask<{T}>({ins[0]} {_CMP[name]} {ins[1]});")
        elif name in _BINFN:
            L.append(f"const {T} {out} = kun::{_BINFN[name]}({ins[0]}, {ins[1]});")
        elif name in _UNARY:
            L.append(f"const {T} {out} = kun::{_UNARY[name]}({ins[0]});")
        elif name == "Log":
            fn = "log_fast" if self.m.options.get("fast_log", True) else "log_lib"
            L.append(f"const {T} {out} = kun::{fn}({ins[0]});")
        elif name == "SetInfOrNanToValue":
            L.append(f"const {T} {out} = kun::set_inf_nan({ins[0]}, {c_literal(op.attrs['value'], T)});")
        elif name == "Select":
            L.append(f"const {T} {out} = kun::sel({ins[0]}, {ins[1]}, {ins[2]});")
        elif name == "IterValue":
            L.append(f"const {T} {out} = {self.win(op.inputs[1], 'iter')};")
        elif name == "WindowLoopIndex":
            w = op.inputs[0].attrs["window"]
            L.append(f"const {T} {out} = ({T})({w} - 1 - iter);")
        elif name == "BackRef":
            L.append(f"const {T} {out} = {self.win(op.inputs[0], op.attrs['window'])};")
        elif name == "FastWindowedSum":
            old = self.win(op.inputs[0], op.attrs["window"])
            L.append(f"const {T} {out} = st{k}.step({old}, {ins[0]});")
        elif name == "ExpMovingAvg":
            w = op.attrs["window"]
            L.append(f"const {T} {out} = st{k}.step({ins[0]}, {T}(2.0) / {T}({w + 1}), "
                     f"{T}(1) - {T}(2.0) / {T}({w + 1}));")
        elif name == "Accumulator":
            L.append(f"const {T} {out} = acc{k};")
        elif name == "SetAccumulator":
            a = self.idx[op.inputs[0]]
            L.append(f"acc{a} = kun::sign_set({ins[1]}) ? {ins[2]} : {ins[0]};")
            L.append(f"const {T} {out} = acc{a};")
        elif name == "WindowedLinearRegression":
            old = self.win(op.inputs[0], op.attrs["window"])
            A = self.AT
            L.append(f"st{k}.step(({A})({old}), ({A})({ins[0]}), ({A}){op.attrs['window']});")
            L.append(f"const {T} {out} = 0;")
        elif name.startswith("WindowedLinearRegression") and name.endswith("Impl"):
            src = self.idx[op.inputs[0]]
            w = op.inputs[0].attrs["window"]
            if name == "WindowedLinearRegressionRSqaureImpl":
                L.append(f"const {T} {out} = ({T})st{src}.rsquare({w});")
            elif name == "WindowedLinearRegressionSlopeImpl":
                L.append(f"const {T} {out} = ({T})st{src}.slope({w});")
            else:
                L.append(f"const {T} {out} = ({T})st{src}.resi({w}, ({self.AT})({ins[1]}));")
        elif name == "SkipListState":
            L.append(f"st{k}.step({ins[0]}, {ins[1]}, (int32_t)({time_idx}), {self.lane_expr(k)});")
            L.append(f"const {T} {out} = 0;")
        elif name in ("SkipListQuantile", "SkipListRank", "SkipListMin", "SkipListMax", "SkipListArgMin"):
            src = self.idx[op.inputs[0]]
            ln = self.lane_expr(src)
            if name == "SkipListQuantile":
                expr = f"st{src}.quantile({c_literal(op.attrs['q'], T)}, {ln})"
            elif name == "SkipListRank":
                expr = f"st{src}.rank({ins[1]}, {ln})"
            elif name == "SkipListMin":
                expr = f"st{src}.vmin_({ln})"
            elif name == "SkipListMax":
                expr = f"st{src}.vmax_({ln})"
            else:
                expr = f"st{src}.argmin((int64_t)({time_idx}), {ln})"
            L.append(f"const {T} {out} = {expr};")
        elif isinstance(op, (ForeachBackWindow, ReductionOp)):
            raise AssertionError("handled by the loop emitter")
        else:
            raise RuntimeError(f"HIP emitter: unsupported op {name}")
        L.extend(self.after_value(k))
        return L

    # None: the emitter has no in-run copy of window loops
    fast_win = None

    def loop_parts(self, loop):
        """(declarations, window, loop body, finalisation, in-run body) of a
        window loop (see loop_code). The in-run body, when the emitter has
        one and it differs, reads row sources without the clamp and NaN
        select that rows before the run start need; it is correct for rows
        i >= window - 1 and reads the same values there."""
        T = self.T
        w = loop.attrs["window"]
        body = [op for op in self.ops if op.get_parent() is loop]
        reds = [op for op in self.ops if isinstance(op, ReductionOp) and op.get_loop() is loop]
        L = []
        for r in reds:
            k = self.idx[r]
            name = r.__class__.__name__
            L.append(f"{T} v{k};")
            if name == "ReduceAdd":
                L.append(f"{self.AT} r{k}_v = 0, r{k}_c = 0;")
            elif name == "ReduceMul":
                L.append(f"{T} r{k}_v = 1;")
            elif name == "ReduceMin":
                L.append(f"{T} r{k}_v = kun::inf<{T}>();")
            elif name == "ReduceMax":
                L.append(f"{T} r{k}_v = -kun::inf<{T}>();")
            elif name in ("ReduceArgMax", "ReduceArgMin"):
                sgn = "-" if name == "ReduceArgMax" else ""
                L.append(f"{T} r{k}_v = {sgn}kun::inf<{T}>(), r{k}_i = 0;")
            elif name == "ReduceRank":
                src = _rank_copy_source(r)
                if src is None:
                    L.append(f"const {T} r{k}_cur = {self.v(r.inputs[1])};")
                    L.append(f"{T} r{k}_lt = 0, r{k}_eq = 0;")
                else:
                    # the reference copy-constructs this rank from `src`
                    # (see _rank_copy_source): it compares against src's
                    # value and continues its counts; src's result is
                    # lt + (eq + 1) / 2, so starting from result - 1/2 and
                    # no equal count gives the same multiples of 1/2 exactly
                    root = src
                    while _rank_copy_source(root) is not None:
                        root = _rank_copy_source(root)
                    have = set(self.ops) | set(getattr(self, "ext", []))
                    if src not in have or root.inputs[1] not in have:
                        raise RuntimeError("HIP emitter: a rank of a rank whose values sit in another kernel")
                    L.append(f"const {T} r{k}_cur = {self.v(root.inputs[1])};")
                    L.append(f"{T} r{k}_lt = {self.v(src)} - {T}(0.5), r{k}_eq = 0;")
            elif name == "ReduceDecayLinear":
                rw = r.attrs["window"]
                L.append(f"const {T} r{k}_step = ({T})(1.0 / ((1.0 + {rw}) * {rw} / 2));")
                L.append(f"{T} r{k}_w = r{k}_step, r{k}_v = 0;")
            else:
                raise RuntimeError(f"HIP emitter: unsupported reduction {name}")
        pre = L
        body_lines = self._loop_body(body, reds)
        fast = None
        if self.fast_win is False:
            # the copy for rows whose window lies inside the run (row >= w - 1)
            self.fast_win = True
            try:
                fast = self._loop_body(body, reds)
            finally:
                self.fast_win = False
            if fast == body_lines:
                fast = None
        L = []
        for r in reds:
            k = self.idx[r]
            name = r.__class__.__name__
            if name == "ReduceAdd" and self.fin:
                L.append(f"v{k} = ({T})r{k}_v;")
            elif name in ("ReduceAdd", "ReduceMul", "ReduceMin", "ReduceMax", "ReduceDecayLinear"):
                L.append(f"v{k} = r{k}_v;")
            elif name in ("ReduceArgMax", "ReduceArgMin"):
                L.append(f"v{k} = r{k}_i;")
            elif name == "ReduceRank":
                L.append(f"v{k} = r{k}_lt + (r{k}_eq + {T}(1.0)) / {T}(2.0);")
            L.extend(self.after_value(k))
        return pre, w, body_lines, L, fast

    def _loop_body(self, body, reds) -> List[str]:
        T = self.T
        L = []
        for op in body:
            L.extend("  " + x for x in self.op_lines(op))
        for r in reds:
            k = self.idx[r]
            x = self.v(r.inputs[0])
            name = r.__class__.__name__
            if name == "ReduceAdd":  # kahanAdd, Ops.hpp:289-298
                A = self.AT
                xv = f"({A}){x}" if self.fin else x
                L.append(f"  {{ const {A} y = {xv} - r{k}_c; const {A} t = r{k}_v + y; "
                         f"r{k}_c = (t - r{k}_v) - y; r{k}_v = t; }}")
            elif name == "ReduceMul":
                L.append(f"  r{k}_v = r{k}_v * {x};")
            elif name in ("ReduceMin", "ReduceMax"):  # Ops.hpp:487-505
                fn = "vmin" if name == "ReduceMin" else "vmax"
                L.append(f"  r{k}_v = (r{k}_v != r{k}_v || {x} != {x}) ? kun::qnan<{T}>() : kun::{fn}(r{k}_v, {x});")
            elif name in ("ReduceArgMax", "ReduceArgMin"):  # Ops.hpp:566-596
                cmp = f"r{k}_v < {x}" if name == "ReduceArgMax" else f"r{k}_v > {x}"
                L.append(f"  {{ const bool nn = (r{k}_v != r{k}_v) || ({x} != {x}); const bool c = {cmp};"
                         f" r{k}_v = c ? {x} : r{k}_v; r{k}_v = nn ? kun::qnan<{T}>() : r{k}_v;"
                         f" r{k}_i = c ? ({T})iter : r{k}_i; r{k}_i = nn ? kun::qnan<{T}>() : r{k}_i; }}")
            elif name == "ReduceRank":  # Ops.hpp:598-616
                L.append(f"  {{ const bool nn = (r{k}_cur != r{k}_cur) || ({x} != {x});"
                         f" r{k}_lt = nn ? kun::qnan<{T}>() : r{k}_lt;"
                         f" r{k}_lt = ({x} < r{k}_cur) ? r{k}_lt + {T}(1.0) : r{k}_lt;"
                         f" r{k}_eq = ({x} == r{k}_cur) ? r{k}_eq + {T}(1.0) : r{k}_eq; }}")
            elif name == "ReduceDecayLinear":  # Ops.hpp:507-520
                L.append(f"  r{k}_v = kun::fma({x}, r{k}_w, r{k}_v); r{k}_w = r{k}_w + r{k}_step;")
        return L

    def step_segments(self, skip=()) -> list:
        """Per top-level op: ("lines", statements) or ("loop", declarations,
        header, body, finalisation)."""
        segs = []
        for op in self.ops:
            if op.get_parent() is not None or isinstance(op, ReductionOp) or op in skip:
                continue
            if isinstance(op, ForeachBackWindow):
                segs.append(("loop",) + tuple(self.loop_parts(op)))
            else:
                segs.append(("lines", self.op_lines(op)))
        return segs

    def step_lines(self, skip=()) -> List[str]:
        return _flatten(self.step_segments(skip), self.loop_unroll)

    def buffer_ptrs(self, args: str = "A") -> List[str]:
        T = self.T
        L, seen = [], set()
        for op in self.ops:
            if isinstance(op, (Input, Output)):
                b = self.m.buffer(op.attrs["name"])
                if b.idx in seen:
                    continue
                seen.add(b.idx)
                const = "const " if isinstance(op, Input) else ""
                L.append(f"{const}{T} *__restrict__ b{b.idx} = ({const}{T} *){args}.bufs[{self.slot[b.idx]}];")
        return L


class StageEmitter(_Lowering):
    """Tick and init kernels of one time-series stage of a STREAM module:
    rings and carried state are [slot][S] arrays in HBM."""

    time_idx = "0"
    # window loops read their rings from HBM: 16 loads in flight per group
    # (a full unroll costs more in registers than it saves in latency)
    loop_unroll = 16

    def __init__(self, module: ModuleIR, stage, kname: str):
        assert module.stream_mode
        self._setup(module, stage)
        if self.fin:
            raise RuntimeError("option fp64_finalize is not supported in stream mode")
        self.kname = kname
        self.ops: List[OpBase] = stage.func.ops
        self.ring: Dict[int, Tuple[int, int]] = {}      # op -> (ring slots, first T slot)
        self.tslot: Dict[int, int] = {}                 # op -> first T slot of its state
        self.islot: Dict[int, int] = {}                 # op -> first int slot of its state
        self.n_t = 0
        self.n_i = 0
        maxoff: Dict[int, int] = {}
        for op in self.ops:
            wr = _window_reads(op)
            if wr:
                k = self.idx[wr[0]]
                maxoff[k] = max(maxoff.get(k, 0), wr[1])
        # a windowed temporary keeps its whole window even where no op of
        # the stage reads that far: the rows are part of the serialised
        # state (the reference stores them in a module buffer of that window)
        for op in self.ops:
            if op.__class__.__name__ == "WindowedTempOutput" and op.get_parent() is None:
                k = self.idx[op]
                maxoff[k] = max(maxoff.get(k, 0), op.attrs["window"] - 1)
        for k in sorted(maxoff):
            slots = _pow2(maxoff[k] + 1)
            self.ring[k] = (slots, self.n_t)
            self.n_t += slots
        for op in self.ops:
            k = self.idx[op]
            name = op.__class__.__name__
            if name == "SkipListState":
                w = op.attrs["window"]
                self.tslot[k] = self.n_t
                self.n_t += w + 1
                self.islot[k] = self.n_i
                self.n_i += w + 3
            else:
                nt, ni = {"FastWindowedSum": (3, 1), "ExpMovingAvg": (1, 0), "Accumulator": (1, 0),
                          "WindowedLinearRegression": (6, 1)}.get(name, (0, 0))
                if nt or ni:
                    self.tslot[k] = self.n_t
                    self.islot[k] = self.n_i
                    self.n_t += nt
                    self.n_i += ni

    @property
    def state_bytes(self) -> int:
        """Per-stock HBM bytes of carried state."""
        return self.n_t * self.esz + self.n_i * 4

    def tmem(self, slot, base: int = None) -> str:
        return f"stT[({slot}) * A.S + s]"

    def imem(self, slot) -> str:
        return f"istp[({slot}) * A.S + s]"

    def lane_expr(self, k: int = None) -> str:
        return "s"

    def win(self, src, off) -> str:
        return self.ring_read(self.idx[src], off)

    def input_expr(self, op) -> str:
        return f"b{self.m.buffer(op.attrs['name']).idx}[s]"

    def output_lines(self, op, val: str) -> List[str]:
        return [f"b{self.m.buffer(op.attrs['name']).idx}[s] = {val};"]

    def state_decls(self) -> List[str]:
        T = self.T
        L = []
        for op in self.ops:
            k = self.idx[op]
            name = op.__class__.__name__
            ts, is_ = self.tslot.get(k), self.islot.get(k)
            if name == "FastWindowedSum":
                L.append(f"  kun::FastSum<{T}> st{k};")
                L.append(f"  st{k}.v = {self.tmem(ts)}; st{k}.cadd = {self.tmem(ts + 1)}; "
                         f"st{k}.csub = {self.tmem(ts + 2)}; st{k}.nnan = {self.imem(is_)};")
            elif name == "WindowedLinearRegression":
                L.append(f"  kun::LinReg<{T}> st{k};")
                fields = ["is", "xs", "x2", "ys", "y2", "xy"]
                L.append("  " + " ".join(f"st{k}.{f} = {self.tmem(ts + j)};" for j, f in enumerate(fields))
                         + f" st{k}.nnan = {self.imem(is_)};")
            elif name == "ExpMovingAvg":
                L.append(f"  kun::Ema<{T}> st{k};")
                L.append(f"  st{k}.v = {self.tmem(ts)};")
            elif name == "Accumulator":
                L.append(f"  {T} acc{k} = {self.tmem(ts)};")
            elif name == "SkipListState":
                w = op.attrs["window"]
                L.append(f"  kun::SortedWin<{T}, {w}> st{k}; st{k}.init(stT + (size_t){ts} * A.S, "
                         f"istp + (size_t){is_} * A.S, A.S);")
                L.append(f"  st{k}.n = {self.imem(is_ + w + 1)}; st{k}.last = {self.imem(is_ + w + 2)};")
        return L

    def state_stores(self) -> List[str]:
        L = []
        for op in self.ops:
            k = self.idx[op]
            name = op.__class__.__name__
            ts, is_ = self.tslot.get(k), self.islot.get(k)
            if name == "FastWindowedSum":
                L.append(f"  {self.tmem(ts)} = st{k}.v; {self.tmem(ts + 1)} = st{k}.cadd; "
                         f"{self.tmem(ts + 2)} = st{k}.csub; {self.imem(is_)} = (int32_t)st{k}.nnan;")
            elif name == "WindowedLinearRegression":
                fields = ["is", "xs", "x2", "ys", "y2", "xy"]
                L.append("  " + " ".join(f"{self.tmem(ts + j)} = st{k}.{f};" for j, f in enumerate(fields))
                         + f" {self.imem(is_)} = (int32_t)st{k}.nnan;")
            elif name == "ExpMovingAvg":
                L.append(f"  {self.tmem(ts)} = st{k}.v;")
            elif name == "Accumulator":
                L.append(f"  {self.tmem(ts)} = acc{k};")
            elif name == "SkipListState":
                w = op.attrs["window"]
                L.append(f"  {self.imem(is_ + w + 1)} = st{k}.n; {self.imem(is_ + w + 2)} = st{k}.last;")
        return L

    def stream_kernels(self) -> str:
        """Per-stock tick and init bodies of the stage (device functions of
        (stage arguments, stock s, tick i); A.S is the stride of the state
        arrays), and the single-context kernels wrapping them: one lane per
        stock of [0, A.S)."""
        T = self.T
        tick = [f"__device__ __forceinline__ void {self.kname}_body(const kun::StreamArgs &A, const long s, "
                f"const long i) {{",
                f"  {T} *__restrict__ stT = ({T} *)A.st;",
                f"  int32_t *__restrict__ istp = (int32_t *)(stT + (size_t){self.n_t} * A.S);"]
        tick.extend("  " + x for x in self.buffer_ptrs())
        tick.extend(self.state_decls())
        tick.append("  {")
        tick.extend("    " + x for x in self.step_lines())
        tick.append("  }")
        tick.extend(self.state_stores())
        tick.append("}")
        tick += [f'extern "C" __global__ __launch_bounds__({TS_BLOCK}) void {self.kname}_tick(kun::StreamArgs A) {{',
                 f"  const long s = (long)blockIdx.x * {TS_BLOCK} + threadIdx.x;",
                 "  if (s >= A.S) return;",
                 f"  {self.kname}_body(A, s, A.tick);",
                 "}"]
        init = [f"__device__ __forceinline__ void {self.kname}_init_body(const kun::StreamArgs &A, const long s) {{",
                f"  {T} *__restrict__ stT = ({T} *)A.st;",
                f"  int32_t *__restrict__ istp = (int32_t *)(stT + (size_t){self.n_t} * A.S);"]
        for k, (slots, base) in sorted(self.ring.items()):
            init.append(f"  for (int q = 0; q < {slots}; ++q) {self.tmem(f'{base} + q')} = kun::qnan<{T}>();")
        for op in self.ops:
            k = self.idx[op]
            name = op.__class__.__name__
            ts, is_ = self.tslot.get(k), self.islot.get(k)
            if name == "FastWindowedSum":
                init.append(f"  {self.tmem(ts)} = 0; {self.tmem(ts + 1)} = 0; {self.tmem(ts + 2)} = 0; "
                            f"{self.imem(is_)} = {op.attrs['window']};")
            elif name == "WindowedLinearRegression":
                init.append("  " + " ".join(f"{self.tmem(ts + j)} = 0;" for j in range(6))
                            + f" {self.imem(is_)} = {op.attrs['window']};")
            elif name == "ExpMovingAvg":
                if len(op.inputs) > 1:
                    raise RuntimeError("EMA with init_val is not supported in stream mode")
                init.append(f"  {self.tmem(ts)} = kun::qnan<{T}>();")
            elif name == "Accumulator":
                init.append(f"  {self.tmem(ts)} = 0;")
            elif name == "SkipListState":
                w = op.attrs["window"]
                init.append(f"  for (int q = 0; q < {w + 1}; ++q) {{ {self.tmem(f'{ts} + q')} = 0; "
                            f"{self.imem(f'{is_} + q')} = 0; }}")
                init.append(f"  {self.imem(is_ + w + 1)} = 0; {self.imem(is_ + w + 2)} = 0;")
        init.append("}")
        init += [f'extern "C" __global__ __launch_bounds__({TS_BLOCK}) void {self.kname}_init(kun::StreamArgs A) {{',
                 f"  const long s = (long)blockIdx.x * {TS_BLOCK} + threadIdx.x;",
                 "  if (s >= A.S) return;",
                 f"  {self.kname}_init_body(A, s);",
                 "}"]
        return "\n".join(tick) + "\n\n" + "\n".join(init)


class CompEmitter(_Lowering):
    """One component of one level of a batch time-series stage.

    Row sources are the values this component reads from HBM: stage inputs
    (Input ops) and values materialised by lower levels. They are loaded
    `prefetch` rows ahead. Window reads of a row source use an LDS ring when
    the budget allows, a prefetched lag load for single-offset reads
    (BackRef, FastWindowedSum, regression sums), or a direct HBM load."""

    def __init__(self, module: ModuleIR, stage, plan: StagePlan, comp: Component, kind: int,
                 lds_budget: Optional[int] = None):
        self._setup(module, stage)
        self.plan = plan
        self.kind = kind
        self.ext = list(comp.ext)
        self.shadow = set(getattr(comp, "shadow", ()))
        self.ops: List[OpBase] = list(comp.ops)
        opset = set(self.ops)
        self.computed = {op for op in self.ops if not isinstance(op, (Input, ConstantOp))}
        self.single = {op for op in self.ops if isinstance(op, Input) and op.attrs.get("single_value", False)}
        srcs = [op for op in self.ops if isinstance(op, Input) and op not in self.single] + self.ext
        self.srcs = sorted(srcs, key=lambda o: self.idx[o])
        srcset = set(self.srcs)
        # window reads
        ring_need: Dict[OpBase, int] = {}   # computed value -> largest offset
        src_iter: Dict[OpBase, int] = {}    # row source read by IterValue -> largest offset
        self.lags: Dict[Tuple[int, int], None] = {}  # (source op idx, offset) read through a lag load
        for op in self.ops:
            wr = _window_reads(op)
            if not wr:
                continue
            src, off = wr
            if src in srcset:
                if op.__class__.__name__ == "IterValue":
                    src_iter[src] = max(src_iter.get(src, 0), off)
                else:
                    self.lags[(self.idx[src], off)] = None
            else:
                assert src in opset, "window read of a value outside the component"
                ring_need[src] = max(ring_need.get(src, 0), off)
        # Row-chunked components with window loops run their rows in
        # interleaved groups of P (see group_lines); P is fixed here because
        # it sizes the rings.
        self.cost = self._cost()
        has_loop = any(isinstance(op, ForeachBackWindow) for op in self.ops)
        self.interleave = 0
        if kind == PAR and has_loop:
            p = next((pp for c, pp in PAR_UNROLL_BY_COST if self.cost >= c), 8)
            p = min(max(p, INTERLEAVE_MIN), INTERLEAVE_MAX)
            p = max(1, min(p, PREFETCH_REGS // max(1, len(self.srcs) + len(self.lags))))
            self.interleave = p if p > 1 else 0
        extra = self.interleave - 1 if self.interleave else 0
        self.ring: Dict[int, Tuple[int, int]] = {}
        self.tslot: Dict[int, int] = {}
        self.islot: Dict[int, int] = {}
        cap = lds_budget_for(module, lds_budget, kind == SEQ) // (self.esz * TS_BLOCK)
        n_lds = 0
        spilled = []
        for op in sorted(ring_need, key=lambda o: self.idx[o]):
            slots = ring_need[op] + 1 + extra
            if n_lds + slots <= cap:
                self.ring[self.idx[op]] = (slots, n_lds)
                n_lds += slots
            else:
                spilled.append((op, slots))
        self.src_ring: Dict[int, int] = {}  # row source idx -> largest offset, ringed in LDS
        self.direct: set = set()            # row sources read by IterValue straight from HBM
        for op in sorted(src_iter, key=lambda o: (-src_iter[o], self.idx[o])):
            slots = src_iter[op] + 1 + extra
            if n_lds + slots <= cap:
                self.ring[self.idx[op]] = (slots, n_lds)
                self.src_ring[self.idx[op]] = src_iter[op]
                n_lds += slots
            else:
                self.direct.add(self.idx[op])
        self.lds_t = n_lds
        n_t = n_lds
        for op, slots in spilled:
            self.ring[self.idx[op]] = (slots, n_t)
            n_t += slots
        n_i = 0
        # sorted windows of a sequential component live in registers while
        # they fit SORTED_REGS values (with row indices where an argmin reads
        # them), else in HBM scratch columns
        self.sorted_regs: Dict[int, bool] = {}
        regs = 0
        for op in self.ops:
            if op.__class__.__name__ == "SkipListState":
                w = op.attrs["window"]
                k = self.idx[op]
                idx = any(o.__class__.__name__ == "SkipListArgMin" and o.inputs[0] is op for o in self.ops)
                need = (w + 1) * (2 if idx else 1)
                if kind == SEQ and regs + need <= SORTED_REGS:
                    regs += need
                    self.sorted_regs[k] = idx
                    continue
                self.tslot[k] = n_t
                n_t += w + 1
                self.islot[k] = n_i
                n_i += w + 1
        self.n_t, self.n_i = n_t, n_i
        self.halo = self._halo() if kind == PAR else 0
        self.emit_stores = True
        nops = len(self.computed)
        # Row loops are unrolled P times; copy q owns prefetch slot q, which
        # always holds the next row r with r % P == q (loaded P rows ahead),
        # so loads stay in flight across P rows of work. Light row bodies
        # need deeper prefetch to cover HBM latency.
        self.depth = 1
        nload = max(1, sum(1 for op in self.srcs if self._src_row_used(op)) + len(self.lags))
        if kind == PAR:
            p = next((pp for c, pp in PAR_UNROLL_BY_COST if self.cost >= c), 8)
        elif has_loop:
            p = SEQ_LOOP_UNROLL
        elif nops > 60:
            p = SEQ_HEAVY_UNROLL
        else:
            p = SEQ_PREFETCH_MAX if nops <= 24 else 4
        self.prefetch = max(1, min(p, PREFETCH_REGS // nload))
        if self.interleave:
            self.prefetch = self.interleave

    # -- planning helpers -------------------------------------------------
    def halo_prefill_rows(self) -> int:
        return max(self.src_ring.values(), default=0)

    def _cost(self) -> int:
        """Estimated VALU instructions per row (loop bodies count once per
        window iteration)."""
        total = 0
        for op in self.ops:
            if isinstance(op, (Input, ConstantOp)):
                continue
            c = _OP_COST.get(op.__class__.__name__, 1)
            loop = op.get_loop() if isinstance(op, ReductionOp) else op.get_parent()
            total += c * (loop.attrs["window"] if loop is not None else 1)
        return max(1, total)

    def _halo(self) -> int:
        """Rows of history a PAR chunk recomputes before its first stored row:
        the longest chain of in-component window offsets."""
        need: Dict[OpBase, int] = {op: 0 for op in self.ops}
        for op in reversed(self.ops):
            if op not in self.computed:
                continue
            n = need[op]
            name = op.__class__.__name__
            for j, inp in enumerate(op.inputs):
                if inp not in self.computed:
                    continue
                if name == "IterValue" and j == 1:
                    add = op.inputs[0].attrs["window"] - 1
                elif name in _LAG_OPS and j == 0:
                    add = op.attrs["window"]
                elif isinstance(op, ForeachBackWindow):
                    add = op.attrs["window"] - 1
                else:
                    add = 0
                need[inp] = max(need[inp], n + add)
        return max((need[op] for op in self.computed), default=0)

    @property
    def lds_bytes(self) -> int:
        return self.lds_t * self.esz * TS_BLOCK

    @property
    def spill_bytes(self) -> int:
        """HBM bytes per (stock, chunk) column: spilled rings and sorted windows."""
        return (self.n_t - self.lds_t) * self.esz + self.n_i * 4

    # -- storage hooks ----------------------------------------------------
    # Rings hold exactly (largest offset + 1) rows; rp<k> is row i's slot,
    # computed once per row on uniform scalars.
    def ring_read(self, k: int, off) -> str:
        slots, base = self.ring[k]
        q = f"(rp{k} - ({off}))"
        return self.tmem(f"{base} + ({q} < 0 ? {q} + {slots} : {q})", base)

    def ring_push(self, k: int) -> List[str]:
        if k not in self.ring:
            return []
        slots, base = self.ring[k]
        return [f"{self.tmem(f'{base} + rp{k}', base)} = v{k};"]

    def ring_positions(self) -> List[str]:
        return [f"const int rp{k} = (int)i % {slots};" for k, (slots, base) in sorted(self.ring.items())]

    def tmem(self, slot, base: int = None) -> str:
        b = base if base is not None else (slot if isinstance(slot, int) else 0)
        if b < self.lds_t:
            return f"lds[(({slot}) << 6) + lane]"
        return f"stT[(size_t)(({slot}) - {self.lds_t}) * cst + col]"

    def imem(self, slot) -> str:
        return f"istp[(size_t)({slot}) * cst + col]"

    def lane_expr(self, k: int = None) -> str:
        # register-resident sorted windows take no column
        return "0" if k in self.sorted_regs else "col"

    def src_load(self, op, row: str) -> str:
        """HBM load of row source `op` at run row `row` (0 <= row < L)."""
        k = self.idx[op]
        if isinstance(op, Input):
            b = self.m.buffer(op.attrs["name"])
            lay = self.buf_layout(op)
            if b.kind == BUF_INPUT:
                return f"b{b.idx}[kun::at<{lay}>(A.start + ({row}), s, A.S, A.T, A.blk)]"
            return f"b{b.idx}[kun::at<{lay}>(({row}), s, A.S, A.L, A.blk)]"
        return f"m{k}[(size_t)({row}) * A.S + s]"

    def mat_buf(self, op) -> int:
        """Entry of kun::StageArgs::bufs holding materialised value `op`
        ([L][S], placed by the runtime's memory plan): after the stage's
        buffers."""
        return len(self.slot) + self.plan.mat[op]

    def lag_load(self, op, row: str) -> str:
        """Row `row` of a row source, NaN before the run start; the load is
        unconditional (row clamped at 0) so the row loop has no branch."""
        return f"kun::nan_before({self.lag_raw(op, row)}, {row})"

    def lag_raw(self, op, row: str) -> str:
        """The load of lag_load without its NaN select (row clamped at 0)."""
        return self.src_load(op, f"(({row}) < 0 ? 0 : ({row}))")

    # set while lowering the copy of a window loop that runs when every row
    # the loop reads lies inside the run (see loop_parts)
    fast_win = False

    def in_run_load(self, op, row: str) -> str:
        """Row `row` (>= 0) of a row source: no clamp, no NaN select. [rows][S]
        buffers load through kun::wld (row offset in a scalar register)."""
        k = self.idx[op]
        e = self.esz
        if isinstance(op, Input):
            b = self.m.buffer(op.attrs["name"])
            if self.buf_layout(op) != "1":
                return self.src_load(op, row)
            r = f"A.start + ({row})" if b.kind == BUF_INPUT else f"({row})"
            return f"kun::wld(b{b.idx}, so, ({r}) * A.S * {e})"
        return f"kun::wld(m{k}, so, ({row}) * A.S * {e})"

    # ring reads of earlier rows hoisted to the top of the row being
    # emitted (see row_segments): None when not collecting
    _hoist: Optional[Dict[str, str]] = None
    _hoist_min = 1

    def win(self, src, off) -> str:
        k = self.idx[src]
        if src in self.computed:
            if self._hoist is not None and isinstance(off, int) and off >= self._hoist_min:
                name = f"lr{k}_{off}"
                self._hoist.setdefault(name, f"const {self.T} {name} = {self.ring_read(k, off)};")
                return name
            return self.ring_read(k, off)
        if off != "iter":
            # single-offset read of a row source: its prefetched lag load (a
            # ring kept for the source's window loops may be shorter)
            return f"lg{k}_{off}"
        if k in self.src_ring:
            return self.ring_read(k, off)
        if self.fast_win:
            return self.in_run_load(src, f"i - ({off})")
        return self.lag_load(src, f"i - ({off})")

    def input_expr(self, op) -> str:
        raise AssertionError("row sources are loaded at the top of the row")

    def output_lines(self, op, val: str) -> List[str]:
        if not self.emit_stores or op in self.shadow:
            return []
        b = self.m.buffer(op.attrs["name"])
        return [f"b{b.idx}[kun::at<{self.buf_layout(op)}>(i, s, A.S, A.L, A.blk)] = {val};"]

    def after_value(self, k: int) -> List[str]:
        L = self.ring_push(k)
        if self.emit_stores and self.st.func.ops[k] in self.plan.mat:
            L.append(f"w{k}[(size_t)i * A.S + s] = v{k};")
        return L

    # -- code -------------------------------------------------------------
    def state_decls(self) -> List[str]:
        T = self.T
        L = []
        for op in self.ops:
            k = self.idx[op]
            name = op.__class__.__name__
            if name == "FastWindowedSum":
                L.append(f"kun::FastSum<{T}> st{k}; st{k}.init({op.attrs['window']});")
            elif name == "WindowedLinearRegression":
                L.append(f"kun::LinReg<{self.AT}> st{k}; st{k}.init({op.attrs['window']});")
            elif name == "ExpMovingAvg":
                L.append(f"kun::Ema<{T}> st{k};")
                if len(op.inputs) > 1:
                    ib = self.m.buffer(op.inputs[1].attrs["name"]).idx
                    L.append(f"st{k}.v = ((const {T} *)A.bufs[{self.slot[ib]}])[s];")
                else:
                    L.append(f"st{k}.v = kun::qnan<{T}>();")
            elif name == "Accumulator":
                L.append(f"{T} acc{k} = 0;")
            elif name == "SkipListState":
                w = op.attrs["window"]
                if k in self.sorted_regs:
                    L.append(f"kun::SortedWinR<{T}, {w}, {'true' if self.sorted_regs[k] else 'false'}> st{k};")
                    continue
                ts, is_ = self.tslot[k], self.islot[k]
                L.append(f"kun::SortedWin<{T}, {w}> st{k}; st{k}.init(stT + (size_t)({ts - self.lds_t}) * cst, "
                         f"istp + (size_t){is_} * cst, cst);")
        return L

    def _src_row_used(self, op) -> bool:
        """True when row source `op` is needed at the current row: read
        elementwise, or pushed into its LDS ring."""
        if self.idx[op] in self.src_ring:
            return True
        for o in self.ops:
            if isinstance(o, ForeachBackWindow):
                continue
            for j, inp in enumerate(o.inputs):
                if inp is not op:
                    continue
                name = o.__class__.__name__
                if (name == "IterValue" and j == 1) or (name == "BackRef" and j == 0):
                    continue
                return True
        return False

    def row_lines(self, q: int, P: int) -> List[str]:
        """Statements of one row `i`. With the row body unrolled P > 1 times,
        prefetch slot q holds this row and is refilled P rows ahead; with
        P == 1 the slots form a shift register `depth` rows deep."""
        return _flatten(self.row_segments(q, P), self.loop_unroll)

    def row_segments(self, q: int, P: int, grouped: bool = False) -> list:
        """row_lines as segments (see step_segments). Single-offset ring
        reads of rows the ring already holds when the row starts (offset >=
        1; >= P when the P rows of a group interleave) are issued at the top
        of the row, before the row's own ring stores, so their LDS latency
        overlaps the row's work instead of its dependent chain."""
        T = self.T
        L = []
        D = P if P > 1 else self.depth
        nxt = f"(i + {D} < c1 ? i + {D} : c1 - 1)"
        L.extend(self.ring_positions())

        # slots are refilled at the end of the row, once the row has used
        # their values: a refill issued while the old value is still live
        # needs another register, and the loop then copies the slots on
        # every iteration, waiting for their loads
        refill: List[str] = []

        def take(name: str, load: str, wrap=lambda x: x) -> str:
            if P > 1:
                refill.append(f"{name}{q} = {load};")
                return f"{wrap(f'{name}{q}')};"
            # one slot (or a shift register): the refill follows the take, so
            # the load overlaps this row's work (at the row's end it would
            # be waited for at the top of the next row)
            moves = "".join(f" {name}{j} = {name}{j + 1};" for j in range(D - 1))
            return f"{wrap(f'{name}0')};{moves} {name}{D - 1} = {load};"

        for op in self.srcs:
            k = self.idx[op]
            if not self._src_row_used(op):
                continue
            L.append(f"const {T} v{k} = " + take(f"pf{k}_", self.src_load(op, nxt)))
            if k in self.src_ring:
                L.extend(self.ring_push(k))
        for (k, off) in self.lags:
            op = self.st.func.ops[k]
            # the slot holds the raw load; the NaN select for rows before
            # the run start is applied here, where the value is used, so
            # the load stays in flight until then
            L.append(f"const {T} lg{k}_{off} = " + take(f"pl{k}_{off}_", self.lag_raw(op, f"{nxt} - {off}"),
                                                          lambda x: f"kun::nan_before({x}, i - {off})"))
        self._hoist, self._hoist_min = {}, (P if grouped else 1)
        try:
            segs = self.step_segments(skip=set(self.srcs) | self.single)
            L.extend(self._hoist.values())
        finally:
            self._hoist = None
        return [("lines", L)] + segs + [("lines", refill)]

    def group_lines(self, P: int) -> List[str]:
        """Rows i0 .. i0 + P - 1 interleaved op by op: each op is emitted for
        the P rows in turn, and a window loop runs once with the P rows'
        bodies side by side, so their dependent chains (Kahan sums, running
        extremes) overlap. Each row computes exactly what it computes alone;
        rings hold P - 1 extra slots because the later rows' pushes precede
        the earlier rows' window reads."""
        keep = {self.idx[op] for op in self.single}
        segs = [self.row_segments(q, P, grouped=True) for q in range(P)]
        L = [f"const long iq{q} = i0 + {q};" for q in range(P)]
        for j, sg in enumerate(segs[0]):
            if sg[0] == "lines":
                for q in range(P):
                    L.extend(_rename_row(segs[q][j][1], q, keep))
                continue
            for q in range(P):
                L.extend(_rename_row(segs[q][j][1], q, keep))
            body, fast = [], []
            for q in range(P):
                body.extend(_rename_row(segs[q][j][3], q, keep))
                if sg[5] is not None:
                    fast.extend(_rename_row(segs[q][j][5], q, keep))
            L.extend(window_loop(sg[2], body, fast if sg[5] is not None else None, self.loop_unroll, "iq0"))
            for q in range(P):
                L.extend(_rename_row(segs[q][j][4], q, keep))
        return L

    def body(self) -> List[str]:
        """Statements of the component, run by one lane for one stock; `e`
        is this workgroup's entry of the launch table, `chunk_id` its chunk."""
        T = self.T
        L = []
        if self.kind == PAR:
            L.append("const long c0 = (long)chunk_id * A.zchunk[e];")
            L.append("const long c1 = c0 + A.zchunk[e] < A.L ? c0 + A.zchunk[e] : A.L;")
            L.append(f"const long rbeg = c0 - {self.halo} > 0 ? c0 - {self.halo} : 0;")
        else:
            L.append("const long c0 = 0, c1 = A.L, rbeg = 0;")
            L.append("(void)c0;")
        # window loops whose rows all lie in the run take their in-run copy
        # (kun::wld) when every byte offset of the [rows][S] buffers fits
        # in 31 bits
        L.append(f"const unsigned so = (unsigned)s * {self.esz}u; (void)so;")
        L.append(f"const bool wok = (A.T > A.L ? A.T : A.L) * A.S * {self.esz} < 2147483647L; (void)wok;")
        if self.n_t > self.lds_t or self.n_i:
            L.append("const long col = (long)chunk_id * A.S + s;")
            L.append("const long cst = A.S * (long)A.znch[e];")
            L.append(f"{T} *__restrict__ stT = ({T} *)((char *)A.scratch + A.zscr[e]);")
            L.append(f"int32_t *__restrict__ istp = (int32_t *)(stT + (size_t){self.n_t - self.lds_t} * cst);")
        L.extend(self.buffer_ptrs())
        for op in self.ext:
            k = self.idx[op]
            L.append(f"const {T} *__restrict__ m{k} = (const {T} *)A.bufs[{self.mat_buf(op)}];")
        for op in self.ops:
            if op in self.computed and op in self.plan.mat:
                k = self.idx[op]
                L.append(f"{T} *__restrict__ w{k} = ({T} *)A.bufs[{self.mat_buf(op)}];")
        for op in sorted(self.single, key=lambda o: self.idx[o]):
            b = self.m.buffer(op.attrs["name"])
            L.append(f"const {T} v{self.idx[op]} = b{b.idx}[s];")
        for k, (slots, base) in sorted(self.ring.items()):
            L.append(f"for (int q = 0; q < {slots}; ++q) {self.tmem(f'{base} + q', base)} = kun::qnan<{T}>();")
        if self.kind == PAR and self.src_ring and self.halo_prefill_rows():
            # history of ringed HBM values before the first computed row:
            # PF rows of every ringed source per step, loads issued before
            # the LDS stores so they overlap
            PF = 8
            mo = max(self.src_ring.values())
            L.append(f"for (long rb = rbeg - {mo} > 0 ? rbeg - {mo} : 0; rb < rbeg; rb += {PF}) {{")
            for k, off in sorted(self.src_ring.items()):
                op = self.st.func.ops[k]
                for q in range(PF):
                    L.append(f"  const {T} pre{k}_{q} = {self.src_load(op, f'(rb + {q} < rbeg ? rb + {q} : rbeg - 1)')};")
            for k, off in sorted(self.src_ring.items()):
                slots, base = self.ring[k]
                for q in range(PF):
                    L.append(f"  if (rb + {q} < rbeg && rb + {q} >= rbeg - {off}) "
                             f"{self.tmem(f'{base} + (int)(rb + {q}) % {slots}', base)} = pre{k}_{q};")
            L.append("}")
        L.extend(self.state_decls())
        P = self.prefetch
        for q in range(P):
            # first row at or after rbeg that belongs to slot q
            row = f"(rbeg + (({q} - rbeg % {P} + {P}) % {P}))" if P > 1 else "rbeg"
            first = f"({row} < c1 ? {row} : c1 - 1)"
            for op in self.srcs:
                if self._src_row_used(op):
                    L.append(f"{T} pf{self.idx[op]}_{q} = {self.src_load(op, first)};")
            for (k, off) in self.lags:
                L.append(f"{T} pl{k}_{off}_{q} = {self.lag_raw(self.st.func.ops[k], f'{first} - {off}')};")
        if self.kind == PAR:
            # halo rows rebuild history and store nothing; stored rows follow
            # in a phase of their own so neither phase branches on the row
            self.emit_stores = False
            L.extend(self._phase("rbeg", "c0", P))
            self.emit_stores = True
            L.extend(self._phase("c0", "c1", P))
        else:
            L.extend(self._phase("0", "c1", P))
        return L

    def _phase(self, lo: str, hi: str, P: int) -> List[str]:
        """Rows [lo, hi) in groups of P aligned to multiples of P (copy q
        handles rows r % P == q): a guarded first group, unguarded full
        groups, a guarded tail."""
        L = []
        if P == 1:
            L.append(f"for (long i = {lo}; i < {hi}; ++i) {{")
            L.extend("  " + x for x in self.row_lines(0, 1))
            L.append("}")
            return L
        L.append("{")
        L.append(f"  long i0 = {lo} - {lo} % {P};")
        if lo != "0":
            for q in range(P):
                L.append(f"  {{ const long i = i0 + {q};")
                L.append(f"  if (i >= {lo} && i < {hi}) {{")
                L.extend("    " + x for x in self.row_lines(q, P))
                L.append("  } }")
            L.append(f"  i0 += {P};")
        L.append(f"  for (; i0 + {P} <= {hi}; i0 += {P}) {{")
        if self.interleave:
            L.extend("    " + x for x in self.group_lines(P))
        else:
            for q in range(P):
                L.append(f"    {{ const long i = i0 + {q};")
                L.extend("      " + x for x in self.row_lines(q, P))
                if ROW_BARRIER is not None:
                    L.append(f"      __builtin_amdgcn_sched_barrier({ROW_BARRIER});")
                L.append("    }")
        L.append("  }")
        for q in range(P - 1):
            L.append(f"  if (i0 + {q} < {hi}) {{ const long i = i0 + {q};")
            L.extend("    " + x for x in self.row_lines(q, P))
            L.append("  }")
        L.append("}")
        return L


class LevelKernel:
    """One kernel of a batch stage: components of one plan level. The launch
    is one flat grid: the runtime's entry table (StageArgs.z*) gives each
    component its own row chunking (balanced by estimated cost, heaviest
    first) and maps a workgroup to (entry, chunk, 64-stock block)."""

    def __init__(self, module: ModuleIR, stage, plan: StagePlan, level, kname: str, lds_budget: int = None,
                 comps=None):
        self.kname = kname
        self.kind = level.kind
        comps = level.comps if comps is None else comps
        if len(comps) > MAX_COMPS:
            raise RuntimeError(f"stage {stage.name}: {len(comps)} components in one kernel (limit {MAX_COMPS})")
        self.comps = [CompEmitter(module, stage, plan, c, level.kind, lds_budget) for c in comps]
        self.T = self.comps[0].T
        self.halo = max(c.halo for c in self.comps)
        self.lds_bytes = max(c.lds_bytes for c in self.comps)
        # algorithmic traffic: each distinct row source read once, each
        # stored value (outputs, materialised values) written once
        io = set()
        for c in self.comps:
            for op in c.srcs:
                io.add(("r", op.attrs["name"]) if isinstance(op, Input) else ("m", c.idx[op]))
            for op in c.ops:
                if isinstance(op, Output) and op not in c.shadow:
                    io.add(("w", op.attrs["name"]))
                elif op in c.computed and op in plan.mat:
                    io.add(("m", c.idx[op]))
        self.io_values = len(io)
        if len(self.comps[0].slot) + len(plan.mat) > MAX_STAGE_BUFS:
            raise RuntimeError(f"stage {stage.name}: {len(self.comps[0].slot)} buffers and {len(plan.mat)} "
                               f"materialised values exceed {MAX_STAGE_BUFS} kernel buffer entries")
        # module buffers read / written, and materialised values produced
        self.reads = sorted({module.buffer(op.attrs["name"]).idx for c in self.comps for op in c.ops
                             if isinstance(op, Input)})
        self.writes = sorted({module.buffer(op.attrs["name"]).idx for c in self.comps for op in c.ops
                              if isinstance(op, Output) and op not in c.shadow})
        self.produced = {op for c in self.comps for op in c.computed if op in plan.mat}
        self.consumed = {op for c in self.comps for op in c.ext}
        self.mat_writes = sorted(plan.mat[op] for op in self.produced)
        self.mat_reads = sorted(plan.mat[op] for op in self.consumed)
        self.kdeps: List[int] = []
        if self.lds_bytes > LDS_LIMIT:
            raise RuntimeError(f"stage {stage.name}: {self.lds_bytes} bytes of LDS exceed the gfx950 limit")

    def source(self) -> str:
        T = self.T
        L = [f'extern "C" __global__ __launch_bounds__({TS_BLOCK}) void {self.kname}(kun::StageArgs A) {{',
             "  const int lane = threadIdx.x;",
             "  const int wg = blockIdx.x;",
             "  int e = 0;",
             "  while (e + 1 < A.nz && wg >= A.zfirst[e + 1]) ++e;",
             "  const int nsb = (int)((A.S + 63) >> 6);",
             "  const int local = wg - A.zfirst[e];",
             "  const int chunk_id = local / nsb;",
             f"  const long s = (long)(local - chunk_id * nsb) * {TS_BLOCK} + lane;"]
        lds_t = max(c.lds_t for c in self.comps)
        if lds_t:
            L.append(f"  __shared__ {T} lds[{lds_t * TS_BLOCK}];")
        L.append("  if (s >= A.S) return;")
        L.append("  switch (A.zcomp[e]) {")
        for z, c in enumerate(self.comps):
            L.append(f"  case {z}: {{")
            L.extend("    " + x for x in c.body())
            L.append("  } break;")
        L.append("  default: break;")
        L.append("  }")
        L.append("}")
        return "\n".join(L)


def stage_plan(module: ModuleIR, stage) -> StagePlan:
    """Kernel plan of a batch stage; KUN_FORCE_SEQ=1 runs every stage as
    one sequential kernel (a debugging aid)."""
    return plan_stage(stage.func.ops, force_seq=os.environ.get("KUN_FORCE_SEQ", "") == "1")


def stream_phases(module: ModuleIR):
    """Launch plan of a STREAM module's session pool: [(kind, stages)] with
    kind 0 -- one kernel in which each stock's lane runs the tick bodies of
    the listed time-series stages in order -- and kind 1 -- the listed
    cross-sectional stages, one workgroup per (stage, session). Phases
    alternate: a time-series stage joins the phase of its time-series
    producers (its lane computes them first) and follows its cross-sectional
    producers; a cross-sectional stage follows all its producers."""
    writer = {}
    for st in module.stages:
        for b in st.out_bufs:
            writer[b.idx] = st
    order, seen = [], set()

    def visit(st):
        if st.idx in seen:
            return
        seen.add(st.idx)
        for b in st.in_bufs:
            w = writer.get(b.idx)
            if w is not None and w is not st:
                visit(w)
        order.append(st)

    for st in module.stages:
        visit(st)
    phase = {}
    for st in order:
        ts = st.kind == STAGE_TS
        p = 0 if ts else 1
        for b in st.in_bufs:
            w = writer.get(b.idx)
            if w is None or w is st:
                continue
            need = phase[w.idx] + (0 if ts and w.kind == STAGE_TS else 1)
            if need % 2 != (0 if ts else 1):
                need += 1
            p = max(p, need)
        phase[st.idx] = p
    phases = []
    for p in range(max(phase.values()) + 1 if phase else 0):
        sts = [st for st in order if phase[st.idx] == p]
        if sts:
            phases.append((p % 2, sts))
    return phases


def emit_module(module: ModuleIR, lds_budget: int = None) -> str:
    """Full HIP translation unit for one module: kernels + descriptor.
    `lds_budget` overrides the LDS bytes per workgroup given to history
    rings before they move to HBM."""
    return emit_module_parts(module, 1, lds_budget)[0]


CS_MAX_BUFS = 16  # KUN_MAX_CS_BUFS


def generic_cs_kernel(module: ModuleIR, stage, kname: str) -> str:
    """Row kernel of a generic cross-sectional stage: the op's head and body
    (C++ statements, MiscOp.GenericCrossSectionalOp) compiled as device
    code, one thread per row. As in the reference (CodegenCpp.py:111-137),
    input_<n> is the stage's n-th input buffer and output_<n> its n-th
    output buffer."""
    op = next(o for o in stage.func.ops if not isinstance(o, (Input, Output)))
    nin, nout = len(stage.in_bufs), len(stage.out_bufs)
    if nin + nout > CS_MAX_BUFS:
        raise RuntimeError(f"{op.__class__.__name__}: more than {CS_MAX_BUFS} buffers in one cross-sectional stage")
    T = "float" if module.dtype == "float" else "double"
    L = [f'extern "C" __global__ __launch_bounds__({TS_BLOCK}) void {kname}_cs(kun::CsArgs A) {{',
         f"  const long __row = (long)blockIdx.x * {TS_BLOCK} + threadIdx.x;",
         "  if (__row >= A.L) return;",
         f"  using T = {T};",
         "  const size_t num_stocks = (size_t)A.S;",
         "  (void)num_stocks;"]
    names = [f"input_{n}" for n in range(nin)] + [f"output_{n}" for n in range(nout)]
    for j, nm in enumerate(names):
        L.append(f"  const kun::CsRow<T> {nm}{{(T *)A.bufs[{j}], A.lay[{j}], A.S, A.rows[{j}], "
                 f"A.row0[{j}] + __row, A.blk}};")
    L.append("  {")
    L.append(op.generate_head())
    L.append("  {")
    L.append(op.generate_body())
    L.append("  }")
    L.append("  }")
    L.append("}")
    return "\n".join(L)


def emit_module_parts(module: ModuleIR, nparts: int, lds_budget: int = None) -> List[str]:
    """The module as up to `nparts` translation units compiled in parallel:
    batch stage kernels spread over the first units by source size, the
    last unit holding the descriptor (with declarations of the kernels'
    host stubs). Stream modules stay one unit (their phase kernels inline
    the stage bodies)."""
    mid = ident(module.name)
    header = ['#include "kun_device.hpp"', '#include "kun_module_abi.h"', ""]
    parts = list(header)
    split = nparts > 1 and not module.stream_mode
    kernel_srcs: List[Tuple[str, str]] = []  # (kernel name, source) of batch kernels
    cs_args = set()                           # those of them taking kun::CsArgs
    meta = []
    stream_ems = {}
    for st in module.stages:
        kname = f"kun_{mid}_s{st.idx}"
        if st.kind == STAGE_GENERIC:
            src = generic_cs_kernel(module, st, kname)
            if split:
                kernel_srcs.append((f"{kname}_cs", src))
                cs_args.add(f"{kname}_cs")
            else:
                parts.append(src)
                parts.append("")
            meta.append((st, [], f"{kname}_cs", None, 0, 0))
        elif st.kind != STAGE_TS:
            meta.append((st, [], None, None, 0, 0))
        elif module.stream_mode:
            em = StageEmitter(module, st, kname)
            stream_ems[st.idx] = em
            parts.append(em.stream_kernels())
            parts.append("")
            meta.append((st, [], f"{kname}_tick", f"{kname}_init", em.state_bytes, 0))
        else:
            plan = stage_plan(module, st)
            kerns = []
            for level in plan.levels:
                # one launch per LDS class, so that components with small
                # rings are not held to the occupancy of the largest
                classes: Dict[int, list] = {}
                for c in level.comps:
                    lds = CompEmitter(module, st, plan, c, level.kind, lds_budget).lds_bytes
                    if level.kind == PAR:
                        cls = next(j for j, lim in enumerate(LDS_CLASSES) if lds <= lim)
                    else:
                        # sequential levels stay one launch (they are
                        # latency-bound) but for components past 40 KiB,
                        # which would hold every workgroup of it to two per CU
                        cls = 0 if lds <= 40 * 1024 else 1
                    classes.setdefault(cls, []).append(c)
                groups = []
                for cls in sorted(classes):
                    cl = classes[cls]
                    groups.extend(cl[g:g + MAX_COMPS] for g in range(0, len(cl), MAX_COMPS))
                for group in groups:
                    lk = LevelKernel(module, st, plan, level, f"{kname}_k{len(kerns)}", lds_budget, group)
                    if split:
                        kernel_srcs.append((lk.kname, lk.source()))
                    else:
                        parts.append(lk.source())
                        parts.append("")
                    lk.kdeps = sorted({j for j, prev in enumerate(kerns) if prev.produced & lk.consumed})
                    kerns.append(lk)
            meta.append((st, kerns, None, None, 0, len(plan.mat)))
    kinds = {"INPUT": 0, "OUTPUT": 1, "TEMP": 2}
    lay = {"STs": 0, "TS": 1, "STREAM": 2}
    parts.append(f"static const KunBufferDesc kun_{mid}_bufs[] = {{")
    for b in module.buffers:
        parts.append(f'  {{"{b.name}", {kinds[b.kind]}, {b.unreliable}, {b.window}}},')
    parts.append("};")
    stage_kind = {"ts": 0, "rank": 1, "scale": 2, "generic": 3}
    for st, kerns, *_ in meta:
        for tag, arr in (("in", [b.idx for b in st.in_bufs]), ("out", [b.idx for b in st.out_bufs]),
                         ("dep", st.deps)):
            body = ", ".join(str(x) for x in arr) if arr else "-1"
            parts.append(f"static const int kun_{mid}_s{st.idx}_{tag}[] = {{{body}}};")
        if kerns:
            for lk in kerns:
                for tag, vals in (("cost", [c.cost for c in lk.comps]), ("halo", [c.halo for c in lk.comps]),
                                  ("spill", [c.spill_bytes for c in lk.comps]), ("reads", lk.reads),
                                  ("writes", lk.writes), ("kdeps", lk.kdeps), ("mreads", lk.mat_reads),
                                  ("mwrites", lk.mat_writes)):
                    body = ", ".join(str(v) for v in vals) if vals else "-1"
                    parts.append(f"static const int {lk.kname}_{tag}[] = {{{body}}};")
            parts.append(f"static const KunKernelDesc kun_{mid}_s{st.idx}_kerns[] = {{")
            for lk in kerns:
                parts.append(f"  {{(const void *)&{lk.kname}, {lk.kind}, {len(lk.comps)}, {lk.halo}, "
                             f"{lk.lds_bytes}, {lk.io_values}, {lk.kname}_cost, {lk.kname}_halo, {lk.kname}_spill, "
                             f"{lk.kname}_reads, {len(lk.reads)}, {lk.kname}_writes, {len(lk.writes)}, "
                             f"{lk.kname}_kdeps, {len(lk.kdeps)}, {lk.kname}_mreads, {len(lk.mat_reads)}, "
                             f"{lk.kname}_mwrites, {len(lk.mat_writes)}}},")
            parts.append("};")
    pool = "0, nullptr, nullptr"
    if module.stream_mode:
        phases = stream_phases(module)
        stateful = [st for st, _, _, _, sbytes, _ in meta if st.kind == STAGE_TS and sbytes]
        for p, (kind, sts) in enumerate(phases):
            parts.append(f"static const int kun_{mid}_ph{p}_stages[] = {{{', '.join(str(st.idx) for st in sts)}}};")
            if kind == 0:
                parts.append(f'extern "C" __global__ __launch_bounds__({TS_BLOCK}) void kun_{mid}_ph{p}'
                             f"(kun::PoolArgs P) {{")
                parts.append("  const kun::PoolBlock e = P.blocks[blockIdx.x];")
                parts.append("  const long s = (long)e.s0 + threadIdx.x;")
                parts.append("  if (s >= e.s1) return;")
                for st in sts:
                    parts.append(f"  kun_{mid}_s{st.idx}_body(P.stages[{st.idx}], s, e.tick);")
                parts.append("}")
        parts.append(f'extern "C" __global__ __launch_bounds__({TS_BLOCK}) void kun_{mid}_pinit(kun::PoolArgs P) {{')
        parts.append("  const kun::PoolBlock e = P.blocks[blockIdx.x];")
        parts.append("  const long s = (long)e.s0 + threadIdx.x;")
        parts.append("  if (s >= e.s1) return;")
        for st in stateful:
            parts.append(f"  kun_{mid}_s{st.idx}_init_body(P.stages[{st.idx}], s);")
        parts.append("}")
        parts.append(f"static const KunPhaseDesc kun_{mid}_phases[] = {{")
        for p, (kind, sts) in enumerate(phases):
            kern = f"(const void *)&kun_{mid}_ph{p}" if kind == 0 else "nullptr"
            parts.append(f"  {{{kind}, {kern}, kun_{mid}_ph{p}_stages, {len(sts)}}},")
        parts.append("};")
        pool = f"{len(phases)}, kun_{mid}_phases, (const void *)&kun_{mid}_pinit"
    wire = "0, nullptr, 0, nullptr, nullptr"
    if module.stream_mode:
        wbufs, wstates = wire_table(module, stream_ems)
        for j, wb in enumerate(wbufs):
            if wb.rings:
                parts.append(f"static const KunRingRef kun_{mid}_w{j}[] = {{"
                             + ", ".join(f"{{{r.stage}, {r.tslot}, {r.slots}}}" for r in wb.rings) + "};")
        parts.append(f"static const KunWireBuf kun_{mid}_wbufs[] = {{")
        for j, wb in enumerate(wbufs):
            rings = f"kun_{mid}_w{j}" if wb.rings else "nullptr"
            parts.append(f'  {{"{wb.name}", {wb.window}, {wb.buf}, {rings}, {len(wb.rings)}}},')
        parts.append("};")
        wst = f"kun_{mid}_wstates" if wstates else "nullptr"
        if wstates:
            parts.append(f"static const KunWireState kun_{mid}_wstates[] = {{")
            for ws in wstates:
                parts.append(f"  {{{ws.kind}, {ws.stage}, {ws.tslot}, {ws.islot}, {ws.window}, "
                             f"{stream_ems[ws.stage].n_t}}},")
            parts.append("};")
        parts.append(f"static const int kun_{mid}_tslots[] = {{"
                     + ", ".join(str(stream_ems[st.idx].n_t if st.idx in stream_ems else 0) for st in module.stages)
                     + "};")
        wire = f"{len(wbufs)}, kun_{mid}_wbufs, {len(wstates)}, {wst}, kun_{mid}_tslots"
    if module.stream_mode:
        parts.append(f"static const void *kun_{mid}_inits[] = {{"
                     + ", ".join(f"(const void *)&{ini}" if ini else "nullptr" for _, _, _, ini, _, _ in meta) + "};")
    parts.append(f"static const KunStageDesc kun_{mid}_stages[] = {{")
    for st, kerns, tname, ini, sbytes, nmat in meta:
        tk = f"(const void *)&{tname}" if tname else "nullptr"
        karr = f"kun_{mid}_s{st.idx}_kerns" if kerns else "nullptr"
        parts.append(f'  {{"{st.name}", {stage_kind[st.kind]}, {karr}, {len(kerns)}, {nmat}, {tk}, {TS_BLOCK},'
                     f" {sbytes}, kun_{mid}_s{st.idx}_in, {len(st.in_bufs)}, kun_{mid}_s{st.idx}_out,"
                     f" {len(st.out_bufs)}, kun_{mid}_s{st.idx}_dep, {len(st.deps)}}},")
    parts.append("};")
    prog = program_text(module).replace("\\", "\\\\").replace('"', '\\"').replace("\n", '\\n"\n"')
    parts.append(f'static const char kun_{mid}_program[] = "{prog}";')
    inits = f"kun_{mid}_inits" if module.stream_mode else "nullptr"
    parts.append(f'extern "C" __attribute__((visibility("default"))) const KunModuleDesc kun_module_{mid} = {{')
    parts.append(f"  KUN_MODULE_ABI_VERSION, \"{module.name}\", {0 if module.dtype == 'float' else 1},"
                 f" {lay[module.input_layout]}, {lay[module.output_layout]}, {module.blocking_len},"
                 f" {1 if module.aligned else 0}, {len(module.buffers)}, kun_{mid}_bufs, {len(module.stages)},"
                 f" kun_{mid}_stages, kun_{mid}_program, {inits}, {pool}, {wire}}};")
    if not split:
        return ["\n".join(parts) + "\n"]
    n = min(nparts - 1, len(kernel_srcs)) or 1
    units: List[List[str]] = [list(header) for _ in range(n)]
    sizes = [0] * n
    for name, src in sorted(kernel_srcs, key=lambda ks: -len(ks[1])):
        j = sizes.index(min(sizes))
        units[j] += [src, ""]
        sizes[j] += len(src)
    # host-only unit: the descriptor points at the kernels' host stubs, which
    # the other units define; its device pass has nothing to compile
    decls = [f'extern "C" __global__ void {name}(kun::{"CsArgs" if name in cs_args else "StageArgs"} A);'
             for name, _ in kernel_srcs]
    desc = header + ["#if !defined(__HIP_DEVICE_COMPILE__)"] + decls + [""] + parts[len(header):] + ["#endif"]
    return ["\n".join(u) + "\n" for u in units] + ["\n".join(desc) + "\n"]
