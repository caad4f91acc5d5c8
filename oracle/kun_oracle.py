"""ctypes front of the CPU oracle (oracle/kun_oracle.c).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg, never by the product package.

    prog = OracleProgram(program_text(module_ir))
    outs = prog.run(inputs, num_stocks, total_time, cur_time, length)
"""
import ctypes
import os
import subprocess
from typing import Dict, Optional

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "_build", "libkunoracle.so")

_lib = None


def build(force: bool = False) -> str:
    """Compile the oracle with its Makefile (gcc, -ffp-contract=off)."""
    if force or not os.path.exists(LIB_PATH):
        subprocess.check_call(["make", "-B" if force else "-s", "-C", HERE],
                              stdout=subprocess.DEVNULL)
    return LIB_PATH


def lib():
    global _lib
    if _lib is None:
        build()
        L = ctypes.CDLL(LIB_PATH)
        L.kun_oracle_load.restype = ctypes.c_void_p
        L.kun_oracle_load.argtypes = [ctypes.c_char_p, ctypes.c_char_p, ctypes.c_int]
        L.kun_oracle_free.argtypes = [ctypes.c_void_p]
        L.kun_oracle_run.restype = ctypes.c_int
        L.kun_oracle_run.argtypes = [ctypes.c_void_p, ctypes.POINTER(ctypes.c_void_p), ctypes.c_long,
                                     ctypes.c_long, ctypes.c_long, ctypes.c_long, ctypes.c_int,
                                     ctypes.c_char_p, ctypes.c_int]
        L.kun_oracle_stream_create.restype = ctypes.c_void_p
        L.kun_oracle_stream_create.argtypes = [ctypes.c_void_p, ctypes.c_long]
        L.kun_oracle_stream_row.restype = ctypes.c_void_p
        L.kun_oracle_stream_row.argtypes = [ctypes.c_void_p, ctypes.c_int]
        L.kun_oracle_stream_run.argtypes = [ctypes.c_void_p]
        L.kun_oracle_stream_free.argtypes = [ctypes.c_void_p]
        vp = ctypes.c_void_p
        L.kun_oracle_corrwith.argtypes = [vp, vp, vp, ctypes.c_long, ctypes.c_long, ctypes.c_long, ctypes.c_long,
                                          ctypes.c_int, ctypes.c_int, ctypes.c_int]
        for sfx in ("f32", "f64"):
            getattr(L, f"kun_oracle_aggregate_{sfx}").argtypes = [vp, vp, ctypes.POINTER(vp), ctypes.c_long,
                                                                  ctypes.c_long, ctypes.c_long, ctypes.c_long]
        _lib = L
    return _lib


class OracleProgram:
    def __init__(self, text: str):
        L = lib()
        err = ctypes.create_string_buffer(512)
        self._h = L.kun_oracle_load(text.encode(), err, 512)
        if not self._h:
            raise RuntimeError("oracle: " + err.value.decode())
        self.buffers = []
        self.dtype = np.float32
        self.in_layout = self.out_layout = "TS"
        self.blocking = 8
        layouts = {"0": "STs", "1": "TS", "2": "STREAM"}
        for line in text.splitlines():
            parts = line.split()
            if not parts:
                continue
            if parts[0] == "B":
                self.buffers.append((int(parts[1]), int(parts[2]), parts[5]))
            elif parts[0] == "dtype":
                self.dtype = np.float64 if parts[1] == "double" else np.float32
            elif parts[0] == "layout":
                self.in_layout, self.out_layout = layouts[parts[1]], layouts[parts[2]]
            elif parts[0] == "blocking":
                self.blocking = int(parts[1])

    def __del__(self):
        if getattr(self, "_h", None) and _lib is not None:
            _lib.kun_oracle_free(self._h)
            self._h = None

    def output_names(self):
        return [n for _, k, n in self.buffers if k == 1]

    def run(self, inputs: Dict[str, np.ndarray], num_stocks: int, total_time: int, cur_time: int,
            length: int, nthreads: int = 1, outputs: Optional[Dict[str, np.ndarray]] = None,
            temps: bool = False):
        """Outputs by name; with temps=True also every temporary buffer
        ([length][num_stocks])."""
        L = lib()
        ptrs = (ctypes.c_void_p * len(self.buffers))()
        keep = []
        result = {}
        for bid, kind, name in self.buffers:
            if kind == 2 and temps:
                arr = np.zeros((length, num_stocks), dtype=self.dtype)
                result[name] = arr
                ptrs[bid] = arr.ctypes.data
            if kind == 0:
                arr = np.ascontiguousarray(inputs[name], dtype=self.dtype)
                keep.append(arr)
                ptrs[bid] = arr.ctypes.data
            elif kind == 1:
                if outputs is not None and name in outputs:
                    arr = outputs[name]
                elif self.out_layout == "STs":
                    b = self.blocking
                    arr = np.empty(((num_stocks + b - 1) // b, length, b), dtype=self.dtype)
                else:
                    arr = np.empty((length, num_stocks), dtype=self.dtype)
                result[name] = arr
                ptrs[bid] = arr.ctypes.data
        err = ctypes.create_string_buffer(512)
        rc = L.kun_oracle_run(self._h, ptrs, num_stocks, total_time, cur_time, length, nthreads, err, 512)
        if rc:
            raise RuntimeError("oracle: " + err.value.decode())
        return result


class OracleStream:
    """Tick-by-tick execution of a STREAM module on the oracle."""

    def __init__(self, prog: OracleProgram, num_stocks: int):
        self.prog = prog
        self.S = num_stocks
        self._h = lib().kun_oracle_stream_create(prog._h, num_stocks)
        self._ids = {name: bid for bid, _, name in prog.buffers}

    def __del__(self):
        if getattr(self, "_h", None) and _lib is not None:
            _lib.kun_oracle_stream_free(self._h)
            self._h = None

    def _row(self, name):
        ptr = lib().kun_oracle_stream_row(self._h, self._ids[name])
        ct = ctypes.c_double if self.prog.dtype == np.float64 else ctypes.c_float
        return np.ctypeslib.as_array(ctypes.cast(ptr, ctypes.POINTER(ct)), shape=(self.S,))

    def push(self, name: str, row: np.ndarray):
        self._row(name)[:] = row

    def run(self):
        lib().kun_oracle_stream_run(self._h)

    def get(self, name: str) -> np.ndarray:
        return self._row(name).copy()


AGG_KINDS = ("sum", "min", "max", "first", "last", "count", "mean")


def corr_with(x: np.ndarray, y: np.ndarray, layout: str = "TS", rank_inputs: bool = False, blocking: int = 8,
              start: int = 0, length: Optional[int] = None) -> np.ndarray:
    """Per-row correlation of x with y over the stocks (float32)."""
    if layout == "TS":
        T, S = x.shape
    else:
        nb, T, blk = x.shape
        S, blocking = nb * blk, blk
    L = T - start if length is None else length
    out = np.empty(L, dtype=np.float32)
    x = np.ascontiguousarray(x, dtype=np.float32)
    y = np.ascontiguousarray(y, dtype=np.float32)
    lib().kun_oracle_corrwith(x.ctypes.data, y.ctypes.data, out.ctypes.data, S, T, start, L,
                              1 if layout == "STs" else 0, blocking, 1 if rank_inputs else 0)
    return out


def aggregate(x: np.ndarray, labels: np.ndarray, start: int = 0, length: Optional[int] = None):
    """{kind: [groups][S]} of the rows of x ([T][S]) grouped by runs of equal
    labels."""
    T, S = x.shape
    L = T - start if length is None else length
    lab = labels[start:start + L]
    groups = 1 + int(np.count_nonzero(lab[1:] != lab[:-1])) if L > 0 else 0
    outs = {k: np.empty((groups, S), dtype=x.dtype) for k in AGG_KINDS}
    ptrs = (ctypes.c_void_p * 7)(*[outs[k].ctypes.data for k in AGG_KINDS])
    x = np.ascontiguousarray(x)
    labels = np.ascontiguousarray(labels, dtype=x.dtype)
    fn = lib().kun_oracle_aggregate_f64 if x.dtype == np.float64 else lib().kun_oracle_aggregate_f32
    fn(x.ctypes.data, labels.ctypes.data, ptrs, S, T, start, L)
    return outs
