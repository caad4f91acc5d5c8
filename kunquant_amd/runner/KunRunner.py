"""KunRunner: the Python surface of the reference's pybind11 module
(cpp/Python/PyBinding.cpp:146-597), implemented over the C ABI of
libkunhip.so (include/kun_capi.h) with ctypes.

Same names, argument meaning, shape/dtype validation and error messages as
the reference: runGraph validates exactly like PyBinding.cpp:242-394 and
raises RuntimeError; StreamContext mirrors PyBinding.cpp:508-596."""
import ctypes
import os
from typing import Dict, Optional

import numpy as np

from .. import PACKAGE_DIR

_LIB_PATH = os.path.join(PACKAGE_DIR, "lib", "libkunhip.so")
_lib = None

_LAYOUTS = {0: "STs", 1: "TS", 2: "STREAM"}


def _load():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(_LIB_PATH):
        raise RuntimeError(f"libkunhip.so is not built ({_LIB_PATH}); run __graft_entry__.build() "
                           f"or kunquant_amd.build.build_runtime()")
    L = ctypes.CDLL(_LIB_PATH, mode=ctypes.RTLD_GLOBAL)
    vp, sz, cp, i = ctypes.c_void_p, ctypes.c_size_t, ctypes.c_char_p, ctypes.c_int

    def sig(name, res, *args):
        f = getattr(L, name)
        f.restype = res
        f.argtypes = list(args)

    sig("kunGetLastError", cp)
    sig("kunCreateSingleThreadExecutor", vp)
    sig("kunCreateMultiThreadExecutor", vp, i)
    sig("kunDestoryExecutor", None, vp)
    sig("kunExecutorSynchronize", i, vp)
    sig("kunLoadLibrary", vp, cp)
    sig("kunGetModuleFromLibrary", vp, vp, cp)
    sig("kunUnloadLibrary", None, vp)
    sig("kunCreateBufferNameMap", vp)
    sig("kunDestoryBufferNameMap", None, vp)
    sig("kunSetBufferNameMap", None, vp, cp, vp)
    sig("kunEraseBufferNameMap", None, vp, cp)
    sig("kunRunGraphEx", i, vp, vp, vp, sz, sz, sz, sz)
    sig("kunRunGraphDevice", i, vp, vp, vp, sz, sz, sz, sz, vp)
    for n in ("kunModuleDtype", "kunModuleInputLayout", "kunModuleOutputLayout", "kunModuleBlockingLen",
              "kunModuleAligned", "kunModuleNumBuffers", "kunModuleNumStages"):
        sig(n, i, vp)
    sig("kunModuleBufferName", cp, vp, i)
    for n in ("kunModuleBufferKind", "kunModuleBufferUnreliable", "kunModuleBufferWindow"):
        sig(n, i, vp, i)
    sig("kunModuleProgram", cp, vp)
    sig("kunModuleStageName", cp, vp, i)
    for n in ("kunModuleStageKind", "kunModuleStageNumInputs", "kunModuleStageNumOutputs"):
        sig(n, i, vp, i)
    sig("kunModuleStageInput", i, vp, i, i)
    sig("kunModuleStageOutput", i, vp, i, i)
    sig("kunExecutorEnableProfiling", i, vp, i)
    sig("kunExecutorStageTime", i, vp, vp, i, ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_long))
    sig("kunExecutorKernelTime", i, vp, vp, i, i, ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_long))
    sig("kunExecutorSetChunking", i, vp, ctypes.c_long, ctypes.c_long, ctypes.c_long)
    sig("kunExecutorReleaseTemps", i, vp)
    sig("kunExecutorSetChunkingFor", i, vp, vp, sz, sz, ctypes.c_long, ctypes.c_long, ctypes.c_long)
    szp = ctypes.POINTER(sz)
    sig("kunRunGraphBatchDevice", i, vp, i, ctypes.POINTER(vp), ctypes.POINTER(vp), szp, szp, szp, szp, vp)
    sig("kunCaptureGraphBatch", vp, vp, i, ctypes.POINTER(vp), ctypes.POINTER(vp), szp, szp, szp, szp)
    sig("kunLaunchGraphBatch", i, vp, vp)
    sig("kunDestroyGraphBatch", None, vp)
    sig("kunDescribeGraphBatch", i, i, ctypes.POINTER(vp), szp, szp, szp, szp, i, ctypes.c_char_p, szp)
    sig("kunModuleLoadSerial", ctypes.c_ulonglong, vp)
    sig("kunModuleStageNumKernels", i, vp, i)
    fpp = ctypes.POINTER(vp)
    sig("kunCorrWith", i, vp, i, i, i, fpp, vp, fpp, sz, sz, sz, sz)
    sig("kunCorrWithDevice", i, vp, i, i, i, fpp, vp, fpp, sz, sz, sz, sz, vp)
    sig("kunAggregrate", i, vp, i, fpp, fpp, i, fpp, sz, sz, sz, sz)
    sig("kunAggregrateDevice", i, vp, i, fpp, fpp, i, fpp, sz, sz, sz, sz, sz, vp)
    sig("kunStreamPoolBytesPerStock", sz, vp)
    sig("kunCreateStreamPool", vp, vp, vp, sz, i, i)
    sig("kunDestroyStreamPool", None, vp)
    sig("kunStreamPoolFreeStocks", sz, vp)
    sig("kunStreamPoolAdmit", i, vp, sz, ctypes.POINTER(i))
    sig("kunStreamPoolRelease", i, vp, i)
    sig("kunStreamPoolSubmit", i, vp, i, ctypes.POINTER(i), ctypes.POINTER(vp), ctypes.POINTER(ctypes.c_ulonglong))
    sig("kunStreamPoolDone", i, vp, ctypes.c_ulonglong, i)
    sig("kunStreamPoolRead", i, vp, i, sz, vp)
    sig("kunStreamPoolRowDevice", vp, vp, i, sz)
    sig("kunStreamPoolTicks", ctypes.c_long, vp, i)
    sig("kunStreamPoolSetReadback", i, vp, i, ctypes.POINTER(sz))
    sig("kunStreamPoolResult", vp, vp, ctypes.c_ulonglong, i, sz)
    sig("kunStreamPoolProfile", i, vp, i)
    sig("kunStreamPoolComputed", i, vp, ctypes.c_ulonglong, i)
    sig("kunStreamPoolPhaseTimes", i, vp, ctypes.c_ulonglong, ctypes.POINTER(ctypes.c_float), i)
    sig("kunStreamPoolAdmitState", i, vp, sz, ctypes.c_char_p, sz, ctypes.POINTER(i))
    sig("kunStreamPoolSerializeStates", i, vp, i, ctypes.c_char_p, ctypes.POINTER(sz))
    ip = ctypes.POINTER(ctypes.c_int)
    sig("kunModuleKernelInfo", i, vp, i, i, ip, ip, ip, ip, ip)
    sig("kunModuleDeviceBytes", sz, vp, sz, sz, ip)
    sig("kunCreateStreamEx", i, vp, vp, sz, vp, ctypes.POINTER(vp))
    sig("kunStreamSerializeStates", i, vp, i, vp, ctypes.POINTER(sz))
    sig("kunQueryBufferHandle", sz, vp, cp)
    sig("kunStreamGetCurrentBuffer", vp, vp, sz)
    sig("kunStreamGetCurrentDeviceBuffer", vp, vp, sz)
    sig("kunStreamPushData", None, vp, sz, vp)
    sig("kunStreamRun", None, vp)
    sig("kunStreamTicks", sz, vp)
    sig("kunDestoryStream", None, vp)
    _lib = L
    return L


def _err() -> str:
    return _load().kunGetLastError().decode()


def getRuntimePath() -> str:
    return _LIB_PATH


class Executor:
    """A device executor (one HIP stream on the current device)."""

    def __init__(self, handle):
        self._h = handle

    def __del__(self):
        if getattr(self, "_h", None) and _lib is not None:
            _lib.kunDestoryExecutor(self._h)
            self._h = None

    def synchronize(self):
        if _load().kunExecutorSynchronize(self._h):
            raise RuntimeError(_err())

    def enable_profiling(self, on: bool = True):
        """Record HIP events around every stage kernel of later batch runs."""
        if _load().kunExecutorEnableProfiling(self._h, 1 if on else 0):
            raise RuntimeError(_err())

    def stage_time(self, mod: "Module", stage: int):
        """(accumulated milliseconds of all kernels of a stage, runs)."""
        ms, n = ctypes.c_double(), ctypes.c_long()
        if _load().kunExecutorStageTime(self._h, mod._h, stage, ctypes.byref(ms), ctypes.byref(n)):
            raise RuntimeError(_err())
        return ms.value, n.value

    def kernel_time(self, mod: "Module", stage: int, kern: int):
        """(accumulated milliseconds, launches) of one kernel of a stage."""
        ms, n = ctypes.c_double(), ctypes.c_long()
        if _load().kunExecutorKernelTime(self._h, mod._h, stage, kern, ctypes.byref(ms), ctypes.byref(n)):
            raise RuntimeError(_err())
        return ms.value, n.value

    def set_chunking_for(self, mod: "Module", num_stocks: int, length: int, target_workgroups: int,
                         halo_factor: int, min_rows: int):
        """Chunking for one (module, stocks, rows) only (see autotune)."""
        if _load().kunExecutorSetChunkingFor(self._h, mod._h, num_stocks, length, target_workgroups, halo_factor,
                                             min_rows):
            raise RuntimeError(_err())

    def release_temps(self):
        """Free the temporaries kept for device-resident runs (waits for the
        device)."""
        if _load().kunExecutorReleaseTemps(self._h):
            raise RuntimeError(_err())

    def set_chunking(self, target_workgroups: int = 32768, halo_factor: int = 3, min_rows: int = 64):
        """How row-chunked stage kernels split the rows (results do not
        depend on it)."""
        if _load().kunExecutorSetChunking(self._h, target_workgroups, halo_factor, min_rows):
            raise RuntimeError(_err())


def createSingleThreadExecutor() -> Executor:
    h = _load().kunCreateSingleThreadExecutor()
    if not h:
        raise RuntimeError(_err())
    return Executor(h)


def createMultiThreadExecutor(num_threads: int) -> Executor:
    h = _load().kunCreateMultiThreadExecutor(int(num_threads))
    if not h:
        raise RuntimeError(_err())
    return Executor(h)


class Module:
    def __init__(self, handle, lib: "Library", name: str = None):
        self._h = handle
        self._lib = lib  # keeps the library loaded
        self.name = name
        L = _load()
        self.dtype = np.float64 if L.kunModuleDtype(handle) == 1 else np.float32
        self.input_layout = _LAYOUTS[L.kunModuleInputLayout(handle)]
        self._output_layout = _LAYOUTS[L.kunModuleOutputLayout(handle)]
        self._blocking = L.kunModuleBlockingLen(handle)
        self.aligned = bool(L.kunModuleAligned(handle))
        self.buffers = []
        for b in range(L.kunModuleNumBuffers(handle)):
            self.buffers.append((L.kunModuleBufferName(handle, b).decode(), L.kunModuleBufferKind(handle, b),
                                 L.kunModuleBufferUnreliable(handle, b), L.kunModuleBufferWindow(handle, b)))

    @property
    def output_layout(self) -> str:
        return self._output_layout

    @property
    def blocking_len(self) -> int:
        return self._blocking

    def getOutputNames(self):
        return [n for n, k, _, _ in self.buffers if k == 1]

    def getInputNames(self):
        return [n for n, k, _, _ in self.buffers if k == 0]

    def getOutputUnreliableCount(self) -> Dict[str, int]:
        return {n: u for n, k, u, _ in self.buffers if k == 1}

    def program(self) -> str:
        return _load().kunModuleProgram(self._h).decode()

    def stages(self):
        """[(name, kind, input buffer indices, output buffer indices)]; kind
        0 time-series, 1 Rank, 2 Scale."""
        L = _load()
        res = []
        for i in range(L.kunModuleNumStages(self._h)):
            ins = [L.kunModuleStageInput(self._h, i, j) for j in range(L.kunModuleStageNumInputs(self._h, i))]
            outs = [L.kunModuleStageOutput(self._h, i, j) for j in range(L.kunModuleStageNumOutputs(self._h, i))]
            res.append((L.kunModuleStageName(self._h, i).decode(), L.kunModuleStageKind(self._h, i), ins, outs))
        return res

    def kernels(self, stage: int):
        """[(kind, components, halo, lds bytes, HBM values per cell)] of the
        kernels a time-series stage launches; kind 0 row-chunked, 1
        sequential."""
        L = _load()
        if L.kunModuleStageKind(self._h, stage) != 0:
            return []
        res = []
        for j in range(L.kunModuleStageNumKernels(self._h, stage)):
            v = [ctypes.c_int() for _ in range(5)]
            if L.kunModuleKernelInfo(self._h, stage, j, *[ctypes.byref(x) for x in v]):
                raise RuntimeError(_err())
            res.append(tuple(x.value for x in v))
        return res

    def device_bytes(self, num_stocks: int, length: int):
        """(bytes, slots): device memory a batch run holds besides its
        inputs and outputs, and the [rows][stocks] slots among them."""
        L = _load()
        n = ctypes.c_int()
        b = L.kunModuleDeviceBytes(self._h, num_stocks, length, ctypes.byref(n))
        if b == 0 and _err():
            raise RuntimeError(_err())
        return b, n.value


class Library:
    def __init__(self, handle, path: str):
        self._h = handle
        self.path = path
        self._cleanup = None

    @staticmethod
    def load(filename: str) -> "Library":
        h = _load().kunLoadLibrary(filename.encode())
        if not h:
            raise RuntimeError("Cannot load library")
        return Library(h, filename)

    def setCleanup(self, fn):
        self._cleanup = fn

    def getModule(self, name: str) -> Module:
        h = _load().kunGetModuleFromLibrary(self._h, name.encode())
        if not h:
            raise RuntimeError("Module name not found")
        return Module(h, self, name)

    def __del__(self):
        if getattr(self, "_h", None) and _lib is not None:
            _lib.kunUnloadLibrary(self._h)
            self._h = None
            if self._cleanup:
                self._cleanup()


def _shape_str(shape) -> str:
    return "(" + ", ".join(str(int(x)) for x in shape) + ")"


def _expect(dtype, arr: np.ndarray, name: str, shape):
    if arr.dtype != dtype:
        raise RuntimeError(("Expecting float buffer at " if dtype == np.float32 else "Expecting double buffer at ")
                           + name)
    if arr.ndim != len(shape) or tuple(arr.shape) != tuple(shape):
        raise RuntimeError(f"Bad shape at {name} expected {_shape_str(shape)} but got {_shape_str(arr.shape)}")
    if any(s <= 0 for s in arr.shape):
        raise RuntimeError("Bad dimension number at " + name)
    if not arr.flags.c_contiguous:
        raise RuntimeError("Bad stride at " + name)


def _name_map(bufs: Dict[str, int]):
    L = _load()
    h = L.kunCreateBufferNameMap()
    for k, v in bufs.items():
        L.kunSetBufferNameMap(h, k.encode(), ctypes.c_void_p(v))
    return h


def runGraph(exec: Executor, mod: Module, inputs: dict, cur_time: int, length: int, outputs: dict = None,
             skip_check: bool = False, num_stocks: int = -1) -> dict:
    """Batch execution (PyBinding.cpp:242-394): copies inputs to HBM, runs
    every stage on the device and returns the outputs as numpy arrays."""
    if outputs is None:
        outputs = {}
    dtype = mod.dtype
    simd_len = mod.blocking_len
    known_S = known_T = known_num_stocks = 0
    bufs: Dict[str, int] = {}
    keep = []
    for name, arr in inputs.items():
        if not isinstance(arr, np.ndarray):
            arr = np.asarray(arr)
        keep.append(arr)
        if skip_check:
            if known_S == 0 and not name.startswith("__init"):
                if mod.input_layout == "STs":
                    known_S, known_T = arr.shape[0], arr.shape[1]
                elif mod.input_layout == "TS":
                    known_S, known_T = arr.shape[1] // simd_len, arr.shape[0]
                    known_num_stocks = arr.shape[1]
            bufs[name] = arr.ctypes.data
            continue
        if arr.dtype != dtype:
            raise RuntimeError(("Expecting float buffer at " if dtype == np.float32 else
                                "Expecting double buffer at ") + name)
        if name.startswith("__init"):
            if arr.ndim != 1:
                raise RuntimeError("Bad Init shape at " + name)
            if not known_num_stocks:
                known_num_stocks = arr.shape[0]
            _expect(dtype, arr, name, (known_num_stocks,))
        elif mod.input_layout == "STs":
            if arr.ndim != 3:
                raise RuntimeError("Bad STs shape at " + name)
            if known_S == 0:
                known_S, known_T = arr.shape[0], arr.shape[1]
                if not known_num_stocks:
                    known_num_stocks = known_S * simd_len
            _expect(dtype, arr, name, (known_S, known_T, simd_len))
        elif mod.input_layout == "TS":
            if arr.ndim != 2:
                raise RuntimeError("Bad TS shape at " + name)
            if known_S == 0:
                known_S, known_T = arr.shape[1] // simd_len, arr.shape[0]
                if not known_num_stocks:
                    known_num_stocks = arr.shape[1]
                if mod.aligned and known_num_stocks % simd_len != 0:
                    raise RuntimeError("Bad shape at " + name)
            _expect(dtype, arr, name, (known_T, known_num_stocks))
        else:
            raise RuntimeError("Unknown layout at " + name)
        bufs[name] = arr.ctypes.data
    if num_stocks < 0:
        num_stocks = known_num_stocks
    if not skip_check:
        if length > known_T:
            raise RuntimeError("Bad parameter: length")
        if mod.input_layout == "STs":
            if num_stocks > known_num_stocks or known_num_stocks <= known_num_stocks - simd_len:
                raise RuntimeError("num_stocks does not match the shape of inputs")
        elif num_stocks != known_num_stocks:
            raise RuntimeError("num_stocks does not match the shape of inputs")
    if mod.output_layout == "STs":
        out_shape = (known_S, length, simd_len)
    else:
        out_shape = (length, num_stocks)
    ret = {}
    for name in mod.getOutputNames():
        if name in outputs:
            arr = outputs[name]
            if not skip_check:
                _expect(dtype, arr, name, out_shape)
        else:
            arr = np.empty(out_shape, dtype=dtype)
        bufs[name] = arr.ctypes.data
        ret[name] = arr
    L = _load()
    h = _name_map(bufs)
    try:
        rc = L.kunRunGraphEx(exec._h, mod._h, h, num_stocks, known_T, cur_time, length)
    finally:
        L.kunDestoryBufferNameMap(h)
    if rc:
        raise RuntimeError(_err())
    return ret


_FLOAT_SIMD_LEN, _DOUBLE_SIMD_LEN = 8, 4
AGGREGRATION_KINDS = ("sum", "min", "max", "first", "last", "count", "mean")


def _check_panel(arr, name: str, layout: str, dtype, known: dict, simd_len: int):
    """The reference's checkInput (PyBinding.cpp:84-122): TS [T][S] or STs
    [S/simd][T][simd], all buffers of one call of one shape."""
    if not isinstance(arr, np.ndarray):
        arr = np.asarray(arr)
    if layout == "STs":
        if arr.ndim != 3:
            raise RuntimeError("Bad STs shape at " + name)
        if not known:
            known.update(S=arr.shape[0], T=arr.shape[1], N=arr.shape[0] * simd_len)
        _expect(dtype, arr, name, (known["S"], known["T"], simd_len))
    elif layout == "TS":
        if arr.ndim != 2:
            raise RuntimeError("Bad TS shape at " + name)
        if not known:
            known.update(S=arr.shape[1] // simd_len, N=arr.shape[1])
        known.setdefault("T", arr.shape[0])
        _expect(dtype, arr, name, (known["T"], known["N"]))
    else:
        raise RuntimeError("Unknown layout at " + name)
    return arr


def corrWith(exec: Executor, inputs, corr_with, outs, layout: str = "TS", rank_inputs: bool = False):
    """Per-row correlation of each input with `corr_with` over the stocks
    (PyBinding.cpp:396-446, CorrWith.hpp:11-121), float32: outs[i][t] is the
    correlation of row t of inputs[i] (of its rank if rank_inputs) with row
    t of corr_with."""
    if layout not in ("TS", "STs"):
        raise RuntimeError("Unknown layout" + layout)
    if len(inputs) != len(outs):
        raise RuntimeError("number of inputs and outputs should match")
    known: dict = {}
    target = _check_panel(corr_with, "corr_with", layout, np.float32, known, _FLOAT_SIMD_LEN)
    ins = [_check_panel(a, f"buffer_{j}", layout, np.float32, known, _FLOAT_SIMD_LEN) for j, a in enumerate(inputs)]
    for o in outs:
        _expect(np.float32, o, "", (known["T"],))
    n = len(ins)
    vp = ctypes.c_void_p
    inp = (vp * max(1, n))(*[a.ctypes.data for a in ins])
    outp = (vp * max(1, n))(*[o.ctypes.data for o in outs])
    T = known["T"]
    if _load().kunCorrWith(exec._h, 0 if layout == "STs" else 1, 1 if rank_inputs else 0, n, inp,
                           target.ctypes.data, outp, known["N"], T, 0, T):
        raise RuntimeError(_err())


def aggregrate(exec: Executor, inputs, labels, outs):
    """Per-stock reductions over runs of equal labels (PyBinding.cpp:448-508,
    Aggregration.cpp:11-105): outs[i] maps kinds (sum, min, max, first,
    last, count, mean) to [groups][S] arrays for inputs[i] ([T][S], TS)
    grouped by labels[i] ([T])."""
    if len(inputs) != len(labels) or len(inputs) != len(outs):
        raise RuntimeError("number of inputs, labels and outputs should match")
    if len(inputs) == 0:
        return
    dtype = np.float32 if np.asarray(inputs[0]).dtype == np.float32 else np.float64
    simd_len = _FLOAT_SIMD_LEN if dtype == np.float32 else _DOUBLE_SIMD_LEN
    known: dict = {}
    n = len(inputs)
    vp = ctypes.c_void_p
    inp, labp = (vp * n)(), (vp * n)()
    outp = (vp * (7 * n))()
    keep = []
    T = 0
    for j in range(n):
        a = _check_panel(inputs[j], f"buffer_{j}", "TS", dtype, known, simd_len)
        T = known["T"]
        lab = np.asarray(labels[j])
        _expect(dtype, lab, "label", (T,))
        keep += [a, lab]
        inp[j], labp[j] = a.ctypes.data, lab.ctypes.data
        groups = 1 + int(np.count_nonzero(lab[1:] != lab[:-1])) if T else 0
        known_out: dict = {"S": known["S"], "N": known["N"]}
        for name, arr in outs[j].items():
            if name not in AGGREGRATION_KINDS:
                raise RuntimeError("Unknown aggregration kind: " + name)
            k = AGGREGRATION_KINDS.index(name)
            arr = _check_panel(arr, f"output_{name}{k}", "TS", dtype, known_out, simd_len)
            if arr.shape[0] < groups:
                raise RuntimeError(f"Bad shape at output_{name}{k}: {groups} label groups need {groups} rows")
            keep.append(arr)
            outp[j * 7 + k] = arr.ctypes.data
    if _load().kunAggregrate(exec._h, n, inp, labp, 1 if dtype == np.float64 else 0, outp, known["N"], T, 0, T):
        raise RuntimeError(_err())


def runGraphDevice(exec: Executor, mod: Module, buffers: Dict[str, int], num_stocks: int, total_time: int,
                   cur_time: int, length: int, stream: int = 0):
    """Device-resident execution: `buffers` maps every input and output name
    to a device pointer (e.g. torch.Tensor.data_ptr()) in the module's
    layouts. Enqueues on `stream` (0: the executor's stream) and returns
    without synchronising."""
    L = _load()
    h = _name_map(buffers)
    try:
        rc = L.kunRunGraphDevice(exec._h, mod._h, h, num_stocks, total_time, cur_time, length,
                                 ctypes.c_void_p(stream) if stream else None)
    finally:
        L.kunDestoryBufferNameMap(h)
    if rc:
        raise RuntimeError(_err())


def _job_arrays(jobs):
    """jobs: [(module, {name: device pointer}, num_stocks, total_time, cur_time, length)]"""
    L = _load()
    n = len(jobs)
    maps = [_name_map(j[1]) for j in jobs]
    mods = (ctypes.c_void_p * n)(*[j[0]._h for j in jobs])
    bufs = (ctypes.c_void_p * n)(*maps)
    cols = [(ctypes.c_size_t * n)(*[int(j[k]) for j in jobs]) for k in range(2, 6)]
    return maps, mods, bufs, cols


def runGraphBatchDevice(exec: Executor, jobs, stream: int = 0):
    """Independent device-resident jobs enqueued together; each job is
    (module, {name: device pointer}, num_stocks, total_time, cur_time,
    length). Returns without synchronising."""
    L = _load()
    maps, mods, bufs, cols = _job_arrays(jobs)
    try:
        rc = L.kunRunGraphBatchDevice(exec._h, len(jobs), mods, bufs, *cols,
                                      ctypes.c_void_p(stream) if stream else None)
    finally:
        for h in maps:
            L.kunDestoryBufferNameMap(h)
    if rc:
        raise RuntimeError(_err())


def describeGraphBatch(jobs, lanes: int = 6) -> str:
    """Host-only text description of the HIP graph BatchGraph builds for
    jobs [(module, num_stocks, total_time, cur_time, length)]: one line per
    graph node ("<id> kernel|empty <job> <stage> <kernel> <grid> <block> :
    <deps>") and per launch-DAG node ("dag <job> <node> <stage> <kernel> :
    <deps>")."""
    L = _load()
    n = len(jobs)
    mods = (ctypes.c_void_p * n)(*[j[0]._h for j in jobs])
    cols = [(ctypes.c_size_t * n)(*[int(j[k]) for j in jobs]) for k in range(1, 5)]
    size = ctypes.c_size_t(0)
    if L.kunDescribeGraphBatch(n, mods, *cols, lanes, None, ctypes.byref(size)):
        raise RuntimeError(_err())
    buf = ctypes.create_string_buffer(size.value)
    if L.kunDescribeGraphBatch(n, mods, *cols, lanes, buf, ctypes.byref(size)):
        raise RuntimeError(_err())
    return buf.value.decode()


class BatchGraph:
    """A batch of jobs built once as a HIP graph and replayed with one
    launch (device pointers are fixed when it is built)."""

    def __init__(self, exec: Executor, jobs):
        L = _load()
        self._keep = [exec] + [j[0] for j in jobs]
        maps, mods, bufs, cols = _job_arrays(jobs)
        try:
            self._h = L.kunCaptureGraphBatch(exec._h, len(jobs), mods, bufs, *cols)
        finally:
            for h in maps:
                L.kunDestoryBufferNameMap(h)
        if not self._h:
            raise RuntimeError(_err())

    def launch(self, stream: int = 0):
        if _load().kunLaunchGraphBatch(self._h, ctypes.c_void_p(stream) if stream else None):
            raise RuntimeError(_err())

    def __del__(self):
        if getattr(self, "_h", None) and _lib is not None:
            _lib.kunDestroyGraphBatch(self._h)
            self._h = None


class StreamContext:
    """Tick-by-tick execution (PyBinding.cpp:508-596); state is resident in
    HBM between ticks."""

    def __init__(self, exec: Executor, mod: Module, stocks: int, init=None):
        L = _load()
        self._mod = mod
        self._exec = exec
        self._S = int(stocks)
        out = ctypes.c_void_p()
        if init is None:
            rc = L.kunCreateStreamEx(exec._h, mod._h, self._S, None, ctypes.byref(out))
        else:
            args = _StreamExtraArgs()
            args.version = 1
            if isinstance(init, str):
                args.init_kind = 1
                self._init_keep = init.encode()
                args.init.path = self._init_keep
            elif isinstance(init, (bytes, bytearray)):
                args.init_kind = 2
                self._init_keep = bytes(init)
                args.init.memory.buffer = self._init_keep
                args.init.memory.size = len(self._init_keep)
            else:
                raise RuntimeError("Bad type for init, expecting filename or bytes")
            rc = L.kunCreateStreamEx(exec._h, mod._h, self._S, ctypes.byref(args), ctypes.byref(out))
        if rc:
            raise RuntimeError(_err() or "Failed to create stream")
        self._h = out.value

    def __del__(self):
        if getattr(self, "_h", None) and _lib is not None:
            _lib.kunDestoryStream(self._h)
            self._h = None

    def queryBufferHandle(self, name: str) -> int:
        h = _load().kunQueryBufferHandle(self._h, name.encode())
        if h == ctypes.c_size_t(-1).value:
            raise RuntimeError(_err())
        return h

    def getCurrentBuffer(self, handle: int) -> np.ndarray:
        ptr = _load().kunStreamGetCurrentBuffer(self._h, handle)
        if not ptr:
            raise RuntimeError(_err())
        ct = ctypes.c_double if self._mod.dtype == np.float64 else ctypes.c_float
        return np.ctypeslib.as_array(ctypes.cast(ptr, ctypes.POINTER(ct)), shape=(self._S,))

    def getCurrentDeviceBuffer(self, handle: int) -> int:
        return _load().kunStreamGetCurrentDeviceBuffer(self._h, handle)

    def pushData(self, handle: int, data: np.ndarray):
        dtype = self._mod.dtype
        if not isinstance(data, np.ndarray) or data.dtype != dtype or not data.flags.c_contiguous:
            raise RuntimeError("Bad input type to push, expecting float")
        if data.ndim != 1 or data.shape[0] != self._S:
            raise RuntimeError("Bad dimension for input data to push")
        _load().kunStreamPushData(self._h, handle, data.ctypes.data)

    def run(self):
        _load().kunStreamRun(self._h)
        e = _err()
        if e:
            raise RuntimeError(e)

    def serializeStates(self, fileNameOrNone=None):
        L = _load()
        if isinstance(fileNameOrNone, str):
            size = ctypes.c_size_t(0)
            rc = L.kunStreamSerializeStates(self._h, 1, ctypes.c_char_p(fileNameOrNone.encode()), ctypes.byref(size))
            if rc:
                raise RuntimeError("Failed to serialize states")
            return None
        size = ctypes.c_size_t(0)
        L.kunStreamSerializeStates(self._h, 2, None, ctypes.byref(size))
        buf = ctypes.create_string_buffer(size.value)
        rc = L.kunStreamSerializeStates(self._h, 2, buf, ctypes.byref(size))
        if rc:
            raise RuntimeError("Failed to serialize states")
        return buf.raw[:size.value]


class StreamPool:
    """Many stream sessions of one STREAM module resident in HBM side by
    side (no counterpart in the reference, which runs one StreamContext per
    session). A session takes `num_stocks` lanes of the pool's
    `capacity_stocks` at admit() and returns them at release(); submit()
    ticks a set of idle sessions once each in a fixed number of launches and
    returns a cycle id; per session, results equal those of a StreamContext
    fed the same rows."""

    def __init__(self, exec: Executor, mod: Module, capacity_stocks: int, max_sessions: int = 1024,
                 max_inflight: int = 2):
        L = _load()
        self._exec = exec
        self._mod = mod
        self._h = L.kunCreateStreamPool(exec._h, mod._h, int(capacity_stocks), int(max_sessions), int(max_inflight))
        if not self._h:
            raise RuntimeError(_err())
        self._inputs = mod.getInputNames()
        self._handle = {n: j for j, (n, *_) in enumerate(mod.buffers)}
        self._stocks: Dict[int, int] = {}

    @staticmethod
    def bytes_per_stock(mod: Module) -> int:
        """HBM bytes one session lane holds (rows and carried state)."""
        return _load().kunStreamPoolBytesPerStock(mod._h)

    def __del__(self):
        if getattr(self, "_h", None) and _lib is not None:
            _lib.kunDestroyStreamPool(self._h)
            self._h = None

    def free_stocks(self) -> int:
        return _load().kunStreamPoolFreeStocks(self._h)

    def admit(self, num_stocks: int, state: bytes = None):
        """Session id, or None when no contiguous range of the pool is free.
        state: a serialised stream state (StreamContext.serializeStates of
        this library or the reference's) the session resumes from."""
        sid = ctypes.c_int(-1)
        if state is not None:
            state = bytes(state)
            rc = _load().kunStreamPoolAdmitState(self._h, int(num_stocks), state, len(state), ctypes.byref(sid))
        else:
            rc = _load().kunStreamPoolAdmit(self._h, int(num_stocks), ctypes.byref(sid))
        if rc == 2:
            return None
        if rc:
            raise RuntimeError(_err())
        self._stocks[sid.value] = int(num_stocks)
        return sid.value

    def release(self, sid: int):
        if _load().kunStreamPoolRelease(self._h, sid):
            raise RuntimeError(_err())
        self._stocks.pop(sid, None)

    class Rows:
        """Checked input rows of one tick of one session, in the module's
        input order (build once, submit many times; keeps the arrays)."""
        __slots__ = ("sid", "arrays", "ptrs")

        def __init__(self, pool: "StreamPool", sid: int, data: Dict[str, np.ndarray]):
            S = pool._stocks.get(sid)
            dtype = pool._mod.dtype
            self.sid = sid
            self.arrays = []
            for name in pool._inputs:
                a = data[name]
                if not isinstance(a, np.ndarray) or a.dtype != dtype or not a.flags.c_contiguous or a.shape != (S,):
                    raise RuntimeError(f"Bad input row {name} for session {sid}")
                self.arrays.append(a)
            self.ptrs = tuple(a.ctypes.data for a in self.arrays)

    def rows(self, sid: int, data: Dict[str, np.ndarray]) -> "StreamPool.Rows":
        return StreamPool.Rows(self, sid, data)

    def submit(self, ticks) -> int:
        """ticks: [(session id, {input name: row ndarray [S]})] or
        [StreamPool.Rows]; one tick of each listed session (the rows are
        copied before the call returns). Returns the cycle id."""
        n = len(ticks)
        nin = len(self._inputs)
        sids = (ctypes.c_int * n)()
        flat = []
        for j, t in enumerate(ticks):
            if not isinstance(t, StreamPool.Rows):
                t = StreamPool.Rows(self, t[0], t[1])
            if t.sid not in self._stocks:
                raise RuntimeError(f"no such session {t.sid}")
            sids[j] = t.sid
            flat.extend(t.ptrs)
        rows = (ctypes.c_void_p * (n * nin))(*flat)
        cyc = ctypes.c_ulonglong(0)
        if _load().kunStreamPoolSubmit(self._h, n, sids, rows, ctypes.byref(cyc)):
            raise RuntimeError(_err())
        return cyc.value

    def done(self, cycle: int, wait: bool = False) -> bool:
        r = _load().kunStreamPoolDone(self._h, cycle, 1 if wait else 0)
        if r < 0:
            raise RuntimeError(_err())
        return r == 1

    def read(self, sid: int, name: str) -> np.ndarray:
        out = np.empty(self._stocks[sid], dtype=self._mod.dtype)
        if _load().kunStreamPoolRead(self._h, sid, self._handle[name], out.ctypes.data):
            raise RuntimeError(_err())
        return out

    def row_device(self, sid: int, name: str) -> int:
        p = _load().kunStreamPoolRowDevice(self._h, sid, self._handle[name])
        if not p:
            raise RuntimeError(_err())
        return p

    def ticks(self, sid: int) -> int:
        return _load().kunStreamPoolTicks(self._h, sid)

    def set_readback(self, names):
        """Buffers whose rows every later cycle copies to pinned host memory
        before it reports done (results to host inside the cycle); [] turns
        it off. Waits for cycles in flight."""
        hs = (ctypes.c_size_t * max(1, len(names)))(*[self._handle[n] for n in names])
        if _load().kunStreamPoolSetReadback(self._h, len(names), hs):
            raise RuntimeError(_err())
        self._readback = list(names)

    def result(self, cycle: int, sid: int, name: str, copy: bool = True) -> np.ndarray:
        """A session's row of a readback buffer as cycle `cycle` left it, from
        the cycle's pinned host block (waits for the cycle); available until
        max_inflight more cycles have been submitted. copy=False gives a view
        that a later cycle overwrites."""
        p = _load().kunStreamPoolResult(self._h, cycle, sid, self._handle[name])
        if not p:
            raise RuntimeError(_err())
        ct = ctypes.c_float if self._mod.dtype == np.float32 else ctypes.c_double
        v = np.ctypeslib.as_array(ctypes.cast(p, ctypes.POINTER(ct)), shape=(self._stocks[sid],))
        return v.copy() if copy else v

    def computed(self, cycle: int, wait: bool = False) -> bool:
        """True once the cycle's phases are done (results final in HBM; the
        copy to host of readback buffers may still run)."""
        r = _load().kunStreamPoolComputed(self._h, cycle, 1 if wait else 0)
        if r < 0:
            raise RuntimeError(_err())
        return r == 1

    def profile(self, on: bool = True):
        """Record per-phase events in later cycles (see phase_times)."""
        if _load().kunStreamPoolProfile(self._h, 1 if on else 0):
            raise RuntimeError(_err())

    def phase_times(self, cycle: int) -> np.ndarray:
        """ms of a profiled cycle: [input staging, phase 0 .. phase n-1,
        readback] (waits for the cycle)."""
        buf = (ctypes.c_float * 4096)()
        k = _load().kunStreamPoolPhaseTimes(self._h, cycle, buf, 4096)
        if k < 0:
            raise RuntimeError(_err())
        return np.array(buf[:k], dtype=np.float64)

    def serializeStates(self, sid: int) -> bytes:
        """The session's state in the reference's stream-state byte layout."""
        L = _load()
        size = ctypes.c_size_t(0)
        if L.kunStreamPoolSerializeStates(self._h, sid, None, ctypes.byref(size)):
            raise RuntimeError(_err())
        buf = ctypes.create_string_buffer(max(size.value, 1))
        if L.kunStreamPoolSerializeStates(self._h, sid, buf, ctypes.byref(size)):
            raise RuntimeError(_err())
        return buf.raw[:size.value]


class _Mem(ctypes.Structure):
    _fields_ = [("buffer", ctypes.c_char_p), ("size", ctypes.c_size_t)]


class _Init(ctypes.Union):
    _fields_ = [("path", ctypes.c_char_p), ("memory", _Mem)]


class _StreamExtraArgs(ctypes.Structure):
    _fields_ = [("version", ctypes.c_size_t), ("init_kind", ctypes.c_int), ("init", _Init)]
