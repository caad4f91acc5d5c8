"""Stream-state wire table: where each part of the reference's serialised
stream state lives in this backend's device state.

The reference serialises a stream context as (Runtime.cpp:534-546)
  1. every module buffer, in buffer-table order, as a raw StreamBuffer
     (StreamBuffer.hpp:8-52): `window` rows of roundUp(S, blocking) values,
     row p holding the value of the last tick t with t % window == p (NaN
     before it is first written), then one size_t position per blocking
     group of stocks, the next row to write (ticks % window; 0 for
     window 1);
  2. every state buffer, in code-generation order (CodegenCpp.py:286-296):
     one object per blocking group -- raw structs (Ops.hpp:14-25: Accumulator
     Ops.hpp:301-313, FastWindowedSum 324-356, ExpMovingAvg 358-375,
     WindowedLinearRegression 377-420) or, for skip lists, the custom format
     of Ops/Quantile.hpp:54-78 and SkipList.cpp:299-364.
In stream mode the reference's buffer table also holds one buffer per
windowed temporary of a time-series stage, named <stage>_<op index> and
registered after that stage's inputs and outputs (Driver.py:216-236,
CodegenCpp.py:177-186).

This backend keeps the newest row of each module buffer in HBM and the
history of every value a stage reads through a window in that stage's ring
([slot][S], slot = tick & (slots - 1)); carried op state sits in the
stage's [slot][S] state columns. The table below maps each wire entry to
those places; the runtime (csrc/stream.hip) builds and parses the bytes."""
from dataclasses import dataclass, field
from typing import Dict, List, Tuple

from ..Driver import BUF_TEMP, STAGE_TS, ModuleIR
from ..Op import Input, Output

# state kinds, in the numbering of kun_module_abi.h (KUN_WS_*)
WS_FASTSUM, WS_EMA, WS_LINREG, WS_ACCUM, WS_SKIPLIST = range(5)
STATE_KINDS = {"FastWindowedSum": WS_FASTSUM, "ExpMovingAvg": WS_EMA, "WindowedLinearRegression": WS_LINREG,
               "Accumulator": WS_ACCUM, "SkipListState": WS_SKIPLIST}


@dataclass
class WireRing:
    stage: int   # stage index
    tslot: int   # first T slot of the ring in the stage's state
    slots: int   # ring length (power of two)


@dataclass
class WireBuf:
    name: str
    window: int
    buf: int                      # module buffer holding the newest row; -1 for a stage temporary
    rings: List[WireRing] = field(default_factory=list)

    @property
    def source(self) -> WireRing:
        """The ring serialisation reads (the longest)."""
        return max(self.rings, key=lambda r: r.slots) if self.rings else None


@dataclass
class WireState:
    kind: int
    stage: int
    tslot: int
    islot: int
    window: int


def wire_table(module: ModuleIR, emitters: Dict[int, object]) -> Tuple[List[WireBuf], List[WireState]]:
    """emitters: stage index -> StageEmitter of every time-series stage."""
    assert module.stream_mode
    order: List[WireBuf] = []
    seen = set()

    def module_buf(b):
        if b.name in seen:
            return
        seen.add(b.name)
        rings = []
        for st in module.stages:
            em = emitters.get(st.idx)
            if em is None:
                continue
            # a stage reads a buffer's history through its Input op, or, for
            # a buffer it writes, through its Output op (CodegenCpp.py:161-174)
            for op in st.func.ops:
                if isinstance(op, (Input, Output)) and op.attrs["name"] == b.name:
                    k = em.idx[op]
                    if k in em.ring:
                        slots, base = em.ring[k]
                        rings.append(WireRing(st.idx, base, slots))
        order.append(WireBuf(b.name, b.window, b.idx, rings))

    # Driver.py:177-182: the graph's inputs and outputs first
    for b in module.buffers:
        if b.kind != BUF_TEMP:
            module_buf(b)
    states: List[WireState] = []
    for st in module.stages:
        for op in st.func.ops:
            if isinstance(op, (Input, Output)):
                module_buf(module.buffer(op.attrs["name"]))
        em = emitters.get(st.idx)
        if em is None:
            continue
        for i, op in enumerate(st.func.ops):
            name = op.__class__.__name__
            if name == "WindowedTempOutput" and op.get_parent() is None:
                k = em.idx[op]
                assert k == i
                w = op.attrs["window"]
                if k not in em.ring:
                    raise RuntimeError(f"stream state: stage {st.name} keeps no history of {st.name}_{i}")
                slots, base = em.ring[k]
                order.append(WireBuf(f"{st.name}_{i}", w, -1, [WireRing(st.idx, base, slots)]))
            elif name in STATE_KINDS:
                k = em.idx[op]
                states.append(WireState(STATE_KINDS[name], st.idx, em.tslot[k], em.islot.get(k, 0),
                                        op.attrs.get("window", 0)))
    for wb in order:
        need = wb.window if (wb.buf < 0 or wb.window > 1) else 0
        if need and (wb.source is None or wb.source.slots < need):
            raise RuntimeError(f"stream state: no ring keeps {need} rows of {wb.name}")
    # every ring must be restorable from the wire state
    covered = {(r.stage, r.tslot) for wb in order for r in wb.rings}
    for si, em in emitters.items():
        for k, (slots, base) in em.ring.items():
            if (si, base) not in covered:
                raise RuntimeError(f"stream state: ring of value {k} in stage {module.stages[si].name} "
                                   f"has no counterpart in the reference's state")
    return order, states


def wire_bytes_fixed(module: ModuleIR, wbufs: List[WireBuf], S: int) -> int:
    """Bytes of the buffer section for S stocks (S a multiple of blocking)."""
    esz = 4 if module.dtype == "float" else 8
    B = module.blocking_len
    return sum(esz * S * wb.window + -(-S // B) * 8 for wb in wbufs)


def module_wire_table(module: ModuleIR) -> Tuple[List[WireBuf], List[WireState]]:
    from .hip import StageEmitter
    ems = {st.idx: StageEmitter(module, st, f"s{st.idx}") for st in module.stages if st.kind == STAGE_TS}
    return wire_table(module, ems)


def parse_state(module: ModuleIR, blob: bytes, S: int, table=None):
    """Splits a serialised stream state into its parts, stocks in order:
    ([{"name", "window", "rows": [window, S] as stored, "pos": [blocks]}],
     [{"kind", "window", fields...}], bytes consumed). Fixed-layout states
    give [S] arrays per field; skip lists give "last" [S] and per stock the
    stored (value, index) pairs in rank order."""
    import numpy as np
    wbufs, wstates = table or module_wire_table(module)
    T = np.float32 if module.dtype == "float" else np.float64
    IT = np.int32 if module.dtype == "float" else np.int64
    B = module.blocking_len
    nblk = -(-S // B)
    mv = memoryview(blob)
    off = 0

    def take(dtype, count):
        nonlocal off
        n = np.dtype(dtype).itemsize * count
        if off + n > len(mv):
            raise ValueError("stream state blob too short")
        a = np.frombuffer(mv[off:off + n], dtype=dtype).copy()
        off += n
        return a

    bufs = []
    for wb in wbufs:
        rows = take(T, wb.window * S).reshape(wb.window, S)
        bufs.append({"name": wb.name, "window": wb.window, "rows": rows, "pos": take(np.uint64, nblk)})
    fields = {WS_FASTSUM: (["v", "cadd", "csub"], "nnan"), WS_LINREG: (["i", "x", "x2", "y", "y2", "xy"], "nnan"),
              WS_EMA: (["v"], None), WS_ACCUM: (["v"], None)}
    states = []
    for ws in wstates:
        st = {"kind": ws.kind, "window": ws.window}
        if ws.kind == WS_SKIPLIST:
            st["last"] = np.zeros(S, np.int32)
            st["values"], st["index"] = [None] * S, [None] * S
            for blk in range(nblk):
                st["last"][blk * B:(blk + 1) * B] = take(np.int32, B)
                for lane in range(B):
                    n = int(take(np.int32, 1)[0])
                    if n < 0 or n > ws.window:
                        raise ValueError("bad skip list size")
                    pairs = np.frombuffer(take(np.uint8, 16 * n).tobytes(),
                                          dtype=[("v", np.float64), ("i", np.uint64)])
                    st["values"][blk * B + lane] = pairs["v"].copy()
                    st["index"][blk * B + lane] = pairs["i"].copy()
        else:
            tf, itf = fields[ws.kind]
            cols = {f: np.empty(S, T) for f in tf}
            if itf:
                cols[itf] = np.empty(S, IT)
            for blk in range(nblk):
                sl = slice(blk * B, (blk + 1) * B)
                for f in tf:
                    cols[f][sl] = take(T, B)
                if itf:
                    cols[itf][sl] = take(IT, B)
            st.update(cols)
        states.append(st)
    return bufs, states, off
