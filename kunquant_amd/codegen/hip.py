"""HIP emitter: lowers the stages of a ModuleIR to gfx950 kernels built from
the device templates in csrc/kun_device.hpp.

Batch time-series stage kernel (one workgroup = 64 stocks, one lane per
stock, TS rows coalesced):
  * every value read through a window (IterValue, BackRef, FastWindowedSum,
    WindowedLinearRegression sources) keeps a per-lane history ring in LDS
    laid out [slot][lane] (bank-conflict free), slot count a power of two,
    NaN-initialised -- exactly the reference's "no history before start"
    rule (Ops.hpp:87-92, 182-186);
  * carried state (FastWindowedSum, EMA, accumulators, regression sums,
    sorted windows) lives in registers / LDS for the whole time loop;
  * window loops run oldest -> newest like CodegenCpp.py:242-270.
Stream stage kernel (STREAM modules): the same op lowering for one row;
rings and carried state live in HBM as [slot][S] arrays owned by the stream
context (the reference's StreamBuffer / StateBuffer, Runtime.cpp:384-546).
Cross-sectional stages (Rank / Scale) are runtime library kernels."""
from typing import Dict, List, Tuple

from ..Driver import BUF_INPUT, BUF_OUTPUT, BUF_TEMP, ModuleIR, STAGE_TS
from ..Op import ConstantOp, ForeachBackWindow, Input, OpBase, Output, ReductionOp
from .program import literal, program_text

TS_BLOCK = 64
MAX_STAGE_BUFS = 256
LDS_LIMIT = 160 * 1024
# LDS per 64-stock workgroup before history rings spill to HBM scratch
LDS_BUDGET = 64 * 1024

_BIN = {"Add": "+", "Sub": "-", "Mul": "*", "Div": "/"}
_CMP = {"GreaterThan": ">", "GreaterEqual": ">=", "LessThan": "<", "LessEqual": "<=", "Equals": "=="}
_CONST_BIN = {"AddConst": "+", "SubConst": "-", "MulConst": "*", "DivConst": "/",
              "GreaterThanConst": ">", "LessThanConst": "<"}
_UNARY = {"Sqrt": "vsqrt", "Abs": "vabs", "Sign": "vsign", "Not": "bnot", "Exp": "exp_fast"}
_BINFN = {"Max": "vmax", "Min": "vmin", "Or": "bor", "And": "band"}


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


def ident(name: str) -> str:
    return "".join(c if c.isalnum() else "_" for c in name)


class StageEmitter:
    """Emits the batch kernel and (for STREAM modules) the tick/init kernels
    of one time-series stage."""

    def __init__(self, module: ModuleIR, stage, kname: str):
        self.m = module
        self.st = stage
        self.kname = kname
        self.stream = module.stream_mode
        self.T = "float" if module.dtype == "float" else "double"
        self.esz = 4 if self.T == "float" else 8
        self.ops: List[OpBase] = stage.func.ops
        self.idx: Dict[OpBase, int] = {op: i for i, op in enumerate(self.ops)}
        self.slot: Dict[int, int] = {}
        for b in list(stage.in_bufs) + list(stage.out_bufs):
            if b.idx not in self.slot:
                self.slot[b.idx] = len(self.slot)
        if len(self.slot) > MAX_STAGE_BUFS:
            raise RuntimeError(f"stage {stage.name} uses {len(self.slot)} buffers (limit {MAX_STAGE_BUFS})")
        # storage plan: T slots and int slots; LDS rows in batch, HBM rows in stream
        self.ring: Dict[int, Tuple[int, int]] = {}      # op -> (ring slots, first T slot)
        self.tslot: Dict[int, int] = {}                 # op -> first T slot of its state
        self.islot: Dict[int, int] = {}                 # op -> first int slot of its state
        self.n_t = 0
        self.n_i = 0
        self._plan()
        # batch kernels: the first lds_t T slots (and all int slots when they
        # fit) live in LDS, the rest in per-stage HBM scratch [slot][S]
        if self.stream:
            self.lds_t, self.lds_i = 0, 0
        else:
            # rings first, in slot order, while they fit the LDS budget; the
            # sorted windows (and their indices) always live in HBM scratch
            self.lds_i = 0
            cap = LDS_BUDGET // (self.esz * TS_BLOCK)
            self.lds_t = 0
            for lo, hi in sorted((base, base + slots) for slots, base in self.ring.values()):
                if hi <= cap:
                    self.lds_t = hi
                else:
                    break

    def _plan(self):
        maxoff: Dict[int, int] = {}
        for op in self.ops:
            name = op.__class__.__name__
            if name == "IterValue":
                src, off = op.inputs[1], op.inputs[0].attrs["window"] - 1
            elif name in ("BackRef", "FastWindowedSum", "WindowedLinearRegression"):
                src, off = op.inputs[0], op.attrs["window"]
            else:
                continue
            k = self.idx[src]
            maxoff[k] = max(maxoff.get(k, 0), off)
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
                self.n_i += w + 1 + (2 if self.stream else 0)
            elif self.stream:
                nt, ni = {"FastWindowedSum": (3, 1), "ExpMovingAvg": (1, 0), "Accumulator": (1, 0),
                          "WindowedLinearRegression": (6, 1)}.get(name, (0, 0))
                if nt or ni:
                    self.tslot[k] = self.n_t
                    self.islot[k] = self.n_i
                    self.n_t += nt
                    self.n_i += ni

    @property
    def lds_bytes(self) -> int:
        return 0 if self.stream else (self.lds_t * self.esz + self.lds_i * 4) * TS_BLOCK

    @property
    def state_bytes(self) -> int:
        """Per-stock HBM bytes: carried state (stream) or spilled rings (batch)."""
        if self.stream:
            return self.n_t * self.esz + self.n_i * 4
        return (self.n_t - self.lds_t) * self.esz + (self.n_i - self.lds_i) * 4

    # -- addressing ------------------------------------------------------
    def tmem(self, slot, base: int = None) -> str:
        """T storage row `slot` of this lane; `base` is the first slot of the
        object the row belongs to (objects never straddle LDS and HBM)."""
        if self.stream:
            return f"stT[({slot}) * A.S + s]"
        b = base if base is not None else (slot if isinstance(slot, int) else 0)
        if b < self.lds_t:
            return f"lds[(({slot}) << 6) + lane]"
        return f"stT[(({slot}) - {self.lds_t}) * A.S + s]"

    def imem(self, slot) -> str:
        if self.stream:
            return f"istp[({slot}) * A.S + s]"
        if self.lds_i:
            return f"ldsi[(({slot}) << 6) + lane]"
        return f"istp[({slot}) * A.S + s]"

    def ring_read(self, k: int, off) -> str:
        slots, base = self.ring[k]
        return self.tmem(f"{base} + (((int)i - ({off})) & {slots - 1})", base)

    def ring_push(self, k: int) -> List[str]:
        if k not in self.ring:
            return []
        slots, base = self.ring[k]
        return [f"{self.tmem(f'{base} + ((int)i & {slots - 1})', base)} = v{k};"]

    def v(self, op) -> str:
        return f"v{self.idx[op]}"

    def buf_layout(self, op) -> str:
        b = self.m.buffer(op.attrs["name"])
        if b.kind == BUF_TEMP:
            return "1"
        lay = self.m.input_layout if b.kind == BUF_INPUT else self.m.output_layout
        return "1" if lay == "TS" else "0"

    # -- ops -------------------------------------------------------------
    def op_lines(self, op) -> List[str]:
        T = self.T
        k = self.idx[op]
        name = op.__class__.__name__
        ins = [self.v(i) for i in op.inputs]
        out = f"v{k}"
        L: List[str] = []
        time_idx = "0" if self.stream else "i"
        if isinstance(op, Input):
            b = self.m.buffer(op.attrs["name"])
            p = f"b{b.idx}"
            if op.attrs.get("single_value", False) or self.stream:
                L.append(f"const {T} {out} = {p}[s];")
            elif b.kind == BUF_INPUT:
                L.append(f"const {T} {out} = {p}[kun::at<{self.buf_layout(op)}>(A.start + i, s, A.S, A.T, A.blk)];")
            else:
                L.append(f"const {T} {out} = {p}[kun::at<{self.buf_layout(op)}>(i, s, A.S, A.L, A.blk)];")
        elif isinstance(op, Output):
            b = self.m.buffer(op.attrs["name"])
            if self.stream:
                L.append(f"b{b.idx}[s] = {ins[0]};")
            else:
                L.append(f"b{b.idx}[kun::at<{self.buf_layout(op)}>(i, s, A.S, A.L, A.blk)] = {ins[0]};")
            L.append(f"const {T} {out} = {ins[0]};")
        elif name in ("WindowedTempOutput", "ReturnFirstValue"):
            L.append(f"const {T} {out} = {ins[0]};")
        elif isinstance(op, ConstantOp):
            L.append(f"const {T} {out} = {c_literal(op.attrs['value'], T)};")
        elif name in _CONST_BIN:
            c = c_literal(op.attrs["value"], T)
            x, y = (c, ins[0]) if op.attrs.get("swap", False) else (ins[0], c)
            expr = f"({x} {_CONST_BIN[name]} {y})"
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
            L.append(f"const {T} {out} = {self.ring_read(self.idx[op.inputs[1]], 'iter')};")
        elif name == "WindowLoopIndex":
            w = op.inputs[0].attrs["window"]
            L.append(f"const {T} {out} = ({T})({w} - 1 - iter);")
        elif name == "BackRef":
            L.append(f"const {T} {out} = {self.ring_read(self.idx[op.inputs[0]], op.attrs['window'])};")
        elif name == "FastWindowedSum":
            old = self.ring_read(self.idx[op.inputs[0]], op.attrs["window"])
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
            old = self.ring_read(self.idx[op.inputs[0]], op.attrs["window"])
            L.append(f"st{k}.step({old}, {ins[0]}, ({T}){op.attrs['window']});")
            L.append(f"const {T} {out} = 0;")
        elif name.startswith("WindowedLinearRegression") and name.endswith("Impl"):
            src = self.idx[op.inputs[0]]
            w = op.inputs[0].attrs["window"]
            if name == "WindowedLinearRegressionRSqaureImpl":
                L.append(f"const {T} {out} = st{src}.rsquare({w});")
            elif name == "WindowedLinearRegressionSlopeImpl":
                L.append(f"const {T} {out} = st{src}.slope({w});")
            else:
                L.append(f"const {T} {out} = st{src}.resi({w}, {ins[1]});")
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
        L.extend(self.ring_push(k))
        return L

    def lane_expr(self, k: int = None) -> str:
        """Column of this lane in sorted-window storage of op k."""
        if self.stream or (k is not None and self.tslot.get(k, 0) >= self.lds_t):
            return "s"
        return "lane"

    def loop_lines(self, loop) -> List[str]:
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
                L.append(f"{T} r{k}_v = 0, r{k}_c = 0;")
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
                L.append(f"const {T} r{k}_cur = {self.v(r.inputs[1])};")
                L.append(f"{T} r{k}_lt = 0, r{k}_eq = 0;")
            elif name == "ReduceDecayLinear":
                rw = r.attrs["window"]
                L.append(f"const {T} r{k}_step = ({T})(1.0 / ((1.0 + {rw}) * {rw} / 2));")
                L.append(f"{T} r{k}_w = r{k}_step, r{k}_v = 0;")
            else:
                raise RuntimeError(f"HIP emitter: unsupported reduction {name}")
        L.append("{")
        L.append("#pragma unroll 4")
        L.append(f"for (int iter = {w - 1}; iter >= 0; --iter) {{")
        for op in body:
            L.extend("  " + x for x in self.op_lines(op))
        for r in reds:
            k = self.idx[r]
            x = self.v(r.inputs[0])
            name = r.__class__.__name__
            if name == "ReduceAdd":  # kahanAdd, Ops.hpp:289-298
                L.append(f"  {{ const {T} y = {x} - r{k}_c; const {T} t = r{k}_v + y; "
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
        L.append("}")
        L.append("}")
        for r in reds:
            k = self.idx[r]
            name = r.__class__.__name__
            if name in ("ReduceAdd", "ReduceMul", "ReduceMin", "ReduceMax", "ReduceDecayLinear"):
                L.append(f"v{k} = r{k}_v;")
            elif name in ("ReduceArgMax", "ReduceArgMin"):
                L.append(f"v{k} = r{k}_i;")
            elif name == "ReduceRank":
                L.append(f"v{k} = r{k}_lt + (r{k}_eq + {T}(1.0)) / {T}(2.0);")
            L.extend(self.ring_push(k))
        return L

    def step_lines(self) -> List[str]:
        L = []
        for op in self.ops:
            if op.get_parent() is not None or isinstance(op, ReductionOp):
                continue
            L.extend(self.loop_lines(op) if isinstance(op, ForeachBackWindow) else self.op_lines(op))
        return L

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
                L.append(f"  {const}{T} *__restrict__ b{b.idx} = ({const}{T} *){args}.bufs[{self.slot[b.idx]}];")
        return L

    # state declarations: batch = fresh locals; stream = loaded from HBM
    def state_decls(self) -> List[str]:
        T = self.T
        L = []
        for op in self.ops:
            k = self.idx[op]
            name = op.__class__.__name__
            ts, is_ = self.tslot.get(k), self.islot.get(k)
            if name == "FastWindowedSum":
                L.append(f"  kun::FastSum<{T}> st{k};")
                if self.stream:
                    L.append(f"  st{k}.v = {self.tmem(ts)}; st{k}.cadd = {self.tmem(ts + 1)}; "
                             f"st{k}.csub = {self.tmem(ts + 2)}; st{k}.nnan = {self.imem(is_)};")
                else:
                    L.append(f"  st{k}.init({op.attrs['window']});")
            elif name == "WindowedLinearRegression":
                L.append(f"  kun::LinReg<{T}> st{k};")
                if self.stream:
                    fields = ["is", "xs", "x2", "ys", "y2", "xy"]
                    L.append("  " + " ".join(f"st{k}.{f} = {self.tmem(ts + j)};" for j, f in enumerate(fields))
                             + f" st{k}.nnan = {self.imem(is_)};")
                else:
                    L.append(f"  st{k}.init({op.attrs['window']});")
            elif name == "ExpMovingAvg":
                L.append(f"  kun::Ema<{T}> st{k};")
                if self.stream:
                    L.append(f"  st{k}.v = {self.tmem(ts)};")
                elif len(op.inputs) > 1:
                    ib = self.m.buffer(op.inputs[1].attrs["name"]).idx
                    L.append(f"  st{k}.v = (({T} *)A.bufs[{self.slot[ib]}])[s];")
                else:
                    L.append(f"  st{k}.v = kun::qnan<{T}>();")
            elif name == "Accumulator":
                L.append(f"  {T} acc{k} = {self.tmem(ts) if self.stream else '0'};")
            elif name == "SkipListState":
                w = op.attrs["window"]
                if self.stream:
                    L.append(f"  kun::SortedWin<{T}, {w}> st{k}; st{k}.init(stT + (size_t){ts} * A.S, "
                             f"istp + (size_t){is_} * A.S, A.S);")
                    L.append(f"  st{k}.n = {self.imem(is_ + w + 1)}; st{k}.last = {self.imem(is_ + w + 2)};")
                else:
                    assert ts >= self.lds_t and not self.lds_i
                    L.append(f"  kun::SortedWin<{T}, {w}> st{k}; st{k}.init(stT + (size_t)({ts - self.lds_t}) * A.S, "
                             f"istp + (size_t){is_} * A.S, A.S);")
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

    # -- kernels ---------------------------------------------------------
    def batch_kernel(self) -> str:
        T = self.T
        if self.lds_bytes > LDS_LIMIT:
            raise RuntimeError(f"stage {self.st.name}: {self.lds_bytes} bytes of LDS exceed the gfx950 limit")
        L = [f'extern "C" __global__ __launch_bounds__({TS_BLOCK}) void {self.kname}(kun::StageArgs A) {{',
             "  const int lane = threadIdx.x;",
             f"  const long s = (long)blockIdx.x * {TS_BLOCK} + lane;"]
        if self.lds_t:
            L.append(f"  __shared__ {T} lds[{self.lds_t * TS_BLOCK}];")
        if self.lds_i:
            L.append(f"  __shared__ int32_t ldsi[{self.lds_i * TS_BLOCK}];")
        L.append("  if (s >= A.S) return;")
        if self.n_t > self.lds_t or self.n_i > self.lds_i:
            L.append(f"  {T} *__restrict__ stT = ({T} *)A.scratch;")
            L.append(f"  int32_t *__restrict__ istp = (int32_t *)(stT + (size_t){self.n_t - self.lds_t} * A.S);")
        L.extend(self.buffer_ptrs())
        for k, (slots, base) in sorted(self.ring.items()):
            L.append(f"  for (int q = 0; q < {slots}; ++q) {self.tmem(f'{base} + q', base)} = kun::qnan<{T}>();")
        L.extend(self.state_decls())
        L.append("  for (long i = 0; i < A.L; ++i) {")
        L.extend("    " + x for x in self.step_lines())
        L.append("  }")
        L.append("}")
        return "\n".join(L)

    def stream_kernels(self) -> str:
        T = self.T
        tick = [f'extern "C" __global__ __launch_bounds__({TS_BLOCK}) void {self.kname}_tick(kun::StreamArgs A) {{',
                f"  const long s = (long)blockIdx.x * {TS_BLOCK} + threadIdx.x;",
                "  if (s >= A.S) return;",
                f"  {T} *__restrict__ stT = ({T} *)A.st;",
                f"  int32_t *__restrict__ istp = (int32_t *)(stT + (size_t){self.n_t} * A.S);",
                "  const long i = A.tick;"]
        tick.extend(self.buffer_ptrs())
        tick.extend(self.state_decls())
        tick.append("  {")
        tick.extend("    " + x for x in self.step_lines())
        tick.append("  }")
        tick.extend(self.state_stores())
        tick.append("}")
        init = [f'extern "C" __global__ __launch_bounds__({TS_BLOCK}) void {self.kname}_init(kun::StreamArgs A) {{',
                f"  const long s = (long)blockIdx.x * {TS_BLOCK} + threadIdx.x;",
                "  if (s >= A.S) return;",
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
        return "\n".join(tick) + "\n\n" + "\n".join(init)


def emit_module(module: ModuleIR) -> str:
    """Full HIP translation unit for one module: kernels + descriptor."""
    mid = ident(module.name)
    parts = ['#include "kun_device.hpp"', '#include "kun_module_abi.h"', ""]
    meta = []
    for st in module.stages:
        if st.kind == STAGE_TS:
            kname = f"kun_{mid}_s{st.idx}"
            em = StageEmitter(module, st, kname)
            if module.stream_mode:
                parts.append(em.stream_kernels())
                meta.append((st, None, f"{kname}_tick", f"{kname}_init", 0, em.state_bytes))
            else:
                parts.append(em.batch_kernel())
                meta.append((st, kname, None, None, em.lds_bytes, em.state_bytes))
            parts.append("")
        else:
            meta.append((st, None, None, None, 0, 0))
    kinds = {"INPUT": 0, "OUTPUT": 1, "TEMP": 2}
    lay = {"STs": 0, "TS": 1, "STREAM": 2}
    parts.append(f"static const KunBufferDesc kun_{mid}_bufs[] = {{")
    for b in module.buffers:
        parts.append(f'  {{"{b.name}", {kinds[b.kind]}, {b.unreliable}, {b.window}}},')
    parts.append("};")
    stage_kind = {"ts": 0, "rank": 1, "scale": 2}
    for st, *_ in meta:
        for tag, arr in (("in", [b.idx for b in st.in_bufs]), ("out", [b.idx for b in st.out_bufs]),
                         ("dep", st.deps)):
            body = ", ".join(str(x) for x in arr) if arr else "-1"
            parts.append(f"static const int kun_{mid}_s{st.idx}_{tag}[] = {{{body}}};")
    if module.stream_mode:
        parts.append(f"static const void *kun_{mid}_inits[] = {{"
                     + ", ".join(f"(const void *)&{ini}" if ini else "nullptr" for _, _, _, ini, _, _ in meta) + "};")
    parts.append(f"static const KunStageDesc kun_{mid}_stages[] = {{")
    for st, kname, tname, ini, lds, sbytes in meta:
        kern = f"(const void *)&{kname}" if kname else "nullptr"
        tk = f"(const void *)&{tname}" if tname else "nullptr"
        parts.append(f'  {{"{st.name}", {stage_kind[st.kind]}, {kern}, {tk}, {TS_BLOCK}, {lds}, {sbytes},'
                     f" kun_{mid}_s{st.idx}_in, {len(st.in_bufs)}, kun_{mid}_s{st.idx}_out, {len(st.out_bufs)},"
                     f" kun_{mid}_s{st.idx}_dep, {len(st.deps)}}},")
    parts.append("};")
    prog = program_text(module).replace("\\", "\\\\").replace('"', '\\"').replace("\n", '\\n"\n"')
    parts.append(f'static const char kun_{mid}_program[] = "{prog}";')
    inits = f"kun_{mid}_inits" if module.stream_mode else "nullptr"
    parts.append(f'extern "C" __attribute__((visibility("default"))) const KunModuleDesc kun_module_{mid} = {{')
    parts.append(f"  KUN_MODULE_ABI_VERSION, \"{module.name}\", {0 if module.dtype == 'float' else 1},"
                 f" {lay[module.input_layout]}, {lay[module.output_layout]}, {module.blocking_len},"
                 f" {1 if module.aligned else 0}, {len(module.buffers)}, kun_{mid}_bufs, {len(module.stages)},"
                 f" kun_{mid}_stages, kun_{mid}_program, {inits}}};")
    return "\n".join(parts) + "\n"
