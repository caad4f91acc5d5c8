"""The C ABI by the reference's entry-point names, in the shape of the
reference's C API test (tests/capi/test_c.cpp:13-176), through ctypes on
libkunhip.so:

* batch: kunCreateSingleThreadExecutor, kunLoadLibrary,
  kunGetModuleFromLibrary, kunCreateBufferNameMap / kunSetBufferNameMap,
  kunRunGraph on the STs module "testRuntimeModule" (output = input +
  input * 2, 24 stocks x 10 rows), then the destroy / unload calls;
* stream: kunCreateStream on "alpha_101_stream", kunStreamSerializeStates
  with the size query (NULL buffer, size 0 -> KUN_INIT_ERROR and the
  required size) and then the real dump (KUN_SUCCESS, same size),
  kunQueryBufferHandle / kunStreamPushData / kunStreamRun /
  kunStreamGetCurrentBuffer, kunCreateStreamEx resuming from the dump;
  plus the reference's argument checks of kunStreamSerializeStates
  (CApi.cpp:140-164).

Bar: outputs within the reference test's 1e-5 and, beyond it, bit-exact
against float32 numpy of the same expression; a resumed stream equal bit
for bit to the stream it was dumped from."""
import ctypes
import os

import numpy as np
import pytest

from kunquant_amd.build import RUNTIME_LIB
from kunquant_amd.workloads import catalog

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")

KUN_INIT_NONE, KUN_INIT_FILE, KUN_INIT_MEMORY = 0, 1, 2
KUN_SUCCESS, KUN_INIT_ERROR, KUN_INVALID_ARGUMENT = 0, 1, 2
FIELDS = ("close", "open", "high", "low", "volume", "amount")


class _Memory(ctypes.Structure):
    _fields_ = [("buffer", ctypes.c_char_p), ("size", ctypes.c_size_t)]


class _Init(ctypes.Union):
    _fields_ = [("path", ctypes.c_char_p), ("memory", _Memory)]


class ExtraArgs(ctypes.Structure):
    _fields_ = [("version", ctypes.c_size_t), ("init_kind", ctypes.c_int), ("init", _Init)]


@pytest.fixture(scope="module")
def capi():
    L = ctypes.CDLL(RUNTIME_LIB)
    vp, sz, cp, i = ctypes.c_void_p, ctypes.c_size_t, ctypes.c_char_p, ctypes.c_int
    fp = ctypes.POINTER(ctypes.c_float)
    for name, res, args in [
            ("kunCreateSingleThreadExecutor", vp, []), ("kunCreateMultiThreadExecutor", vp, [i]),
            ("kunDestoryExecutor", None, [vp]), ("kunLoadLibrary", vp, [cp]),
            ("kunGetModuleFromLibrary", vp, [vp, cp]), ("kunUnloadLibrary", None, [vp]),
            ("kunCreateBufferNameMap", vp, []), ("kunDestoryBufferNameMap", None, [vp]),
            ("kunSetBufferNameMap", None, [vp, cp, vp]), ("kunEraseBufferNameMap", None, [vp, cp]),
            ("kunRunGraph", None, [vp, vp, vp, sz, sz, sz, sz]), ("kunCreateStream", vp, [vp, vp, sz]),
            ("kunCreateStreamEx", i, [vp, vp, sz, ctypes.POINTER(ExtraArgs), ctypes.POINTER(vp)]),
            ("kunStreamSerializeStates", i, [vp, i, vp, ctypes.POINTER(sz)]),
            ("kunQueryBufferHandle", sz, [vp, cp]), ("kunStreamGetCurrentBuffer", fp, [vp, sz]),
            ("kunStreamPushData", None, [vp, sz, vp]), ("kunStreamRun", None, [vp]),
            ("kunDestoryStream", None, [vp]), ("kunGetLastError", cp, [])]:
        f = getattr(L, name)
        f.restype, f.argtypes = res, args
    return L


def test_batch_by_reference_names(capi):
    """test_c.cpp testBatch: STs [3][10][8] input, output = input * 3."""
    L = capi
    rng = np.random.default_rng(114514)
    num_stocks, num_time = 24, 10
    inputs = rng.random((num_stocks // 8, num_time, 8), dtype=np.float32)
    outputs = np.full_like(inputs, np.nan)
    ex = L.kunCreateSingleThreadExecutor()
    assert ex
    lib = L.kunLoadLibrary(catalog.library_path("testRuntimeModule").encode())
    assert lib
    modu = L.kunGetModuleFromLibrary(lib, b"testRuntimeModule")
    assert modu
    assert not L.kunGetModuleFromLibrary(lib, b"testRuntimeModule2")
    bufs = L.kunCreateBufferNameMap()
    assert bufs
    L.kunSetBufferNameMap(bufs, b"input", inputs.ctypes.data)
    L.kunSetBufferNameMap(bufs, b"output", outputs.ctypes.data)
    L.kunRunGraph(ex, modu, bufs, num_stocks, num_time, 0, num_time)
    assert L.kunGetLastError() == b""
    assert np.abs(outputs - inputs * 3).max() <= 1e-5
    np.testing.assert_array_equal(outputs, inputs + inputs * np.float32(2))
    # an unknown buffer name is reported, not thrown across the C boundary
    L.kunEraseBufferNameMap(bufs, b"input")
    L.kunRunGraph(ex, modu, bufs, num_stocks, num_time, 0, num_time)
    assert b"input" in L.kunGetLastError()
    L.kunDestoryBufferNameMap(bufs)
    L.kunUnloadLibrary(lib)
    L.kunDestoryExecutor(ex)


def _rows(rng, n, num_stocks):
    return [{k: rng.random(num_stocks, dtype=np.float32) for k in FIELDS} for _ in range(n)]


def _push_run(L, ctx, handles, row, outs):
    for k in FIELDS:
        L.kunStreamPushData(ctx, handles[k], row[k].ctypes.data)
    L.kunStreamRun(ctx)
    res = {}
    for name in outs:
        p = L.kunStreamGetCurrentBuffer(ctx, handles[name])
        assert p
        res[name] = np.ctypeslib.as_array(p, shape=(len(row["close"]),)).copy()
    return res


def _alpha101(row):
    return (row["close"] - row["open"]) / ((row["high"] - row["low"]) + np.float32(0.001))


def test_stream_by_reference_names(capi, tmp_path):
    """test_c.cpp testStream: size query, dump, run, resume from the dump."""
    L = capi
    rng = np.random.default_rng(114514)
    num_stocks = 24
    outs = catalog.module_ir("alpha_101_stream").output_names()
    ex = L.kunCreateSingleThreadExecutor()
    lib = L.kunLoadLibrary(catalog.library_path("alpha_101_stream").encode())
    modu = L.kunGetModuleFromLibrary(lib, b"alpha_101_stream")
    assert ex and lib and modu
    ctx = L.kunCreateStream(ex, modu, num_stocks)
    assert ctx
    handles = {k: L.kunQueryBufferHandle(ctx, k.encode()) for k in FIELDS + tuple(outs)}

    # size query: NULL buffer and size 0 -> KUN_INIT_ERROR with the size
    buf_size = ctypes.c_size_t(0)
    assert L.kunStreamSerializeStates(ctx, KUN_INIT_MEMORY, None, ctypes.byref(buf_size)) == KUN_INIT_ERROR
    assert buf_size.value > 0
    size2 = buf_size.value
    blob0 = ctypes.create_string_buffer(size2)
    assert L.kunStreamSerializeStates(ctx, KUN_INIT_MEMORY, blob0, ctypes.byref(buf_size)) == KUN_SUCCESS
    assert buf_size.value == size2
    # the reference's argument checks
    assert L.kunStreamSerializeStates(ctx, KUN_INIT_MEMORY, blob0, None) == KUN_INVALID_ARGUMENT
    five = ctypes.c_size_t(5)
    assert L.kunStreamSerializeStates(ctx, KUN_INIT_MEMORY, None, ctypes.byref(five)) == KUN_INVALID_ARGUMENT
    small = ctypes.c_size_t(size2 - 1)
    tiny = ctypes.create_string_buffer(size2 - 1)
    assert L.kunStreamSerializeStates(ctx, KUN_INIT_MEMORY, tiny, ctypes.byref(small)) == KUN_INIT_ERROR
    assert small.value == size2
    assert L.kunStreamSerializeStates(ctx, 7, blob0, ctypes.byref(buf_size)) == KUN_INVALID_ARGUMENT

    rows = _rows(rng, 14, num_stocks)
    first = _push_run(L, ctx, handles, rows[0], outs)
    assert np.abs(first["alpha101"] - _alpha101(rows[0])).max() <= 1e-5
    np.testing.assert_array_equal(first["alpha101"], _alpha101(rows[0]))
    L.kunDestoryStream(ctx)

    # resume from the dump taken before any tick: same first tick
    extra = ExtraArgs(version=1, init_kind=KUN_INIT_MEMORY)
    extra.init.memory.buffer = ctypes.cast(blob0, ctypes.c_char_p)
    extra.init.memory.size = size2
    ctx = ctypes.c_void_p()
    assert L.kunCreateStreamEx(ex, modu, num_stocks, ctypes.byref(extra), ctypes.byref(ctx)) == KUN_SUCCESS
    assert ctx.value
    again = _push_run(L, ctx, handles, rows[0], outs)
    for k in outs:
        np.testing.assert_array_equal(again[k], first[k], err_msg=k)

    # a dump after 12 ticks: a context resumed from it (memory and file) and
    # the original give the same next ticks, every output bit for bit
    for r in rows[1:12]:
        _push_run(L, ctx, handles, r, outs)
    n = ctypes.c_size_t(0)
    assert L.kunStreamSerializeStates(ctx, KUN_INIT_MEMORY, None, ctypes.byref(n)) == KUN_INIT_ERROR
    blob = ctypes.create_string_buffer(n.value)
    assert L.kunStreamSerializeStates(ctx, KUN_INIT_MEMORY, blob, ctypes.byref(n)) == KUN_SUCCESS
    path = str(tmp_path / "states.bin")
    assert L.kunStreamSerializeStates(ctx, KUN_INIT_FILE, path.encode(), None) == KUN_SUCCESS
    assert open(path, "rb").read() == blob.raw[:n.value]
    mem = ExtraArgs(version=1, init_kind=KUN_INIT_MEMORY)
    mem.init.memory.buffer = ctypes.cast(blob, ctypes.c_char_p)
    mem.init.memory.size = n.value
    fil = ExtraArgs(version=1, init_kind=KUN_INIT_FILE)
    fil.init.path = path.encode()
    resumed = []
    for extra in (mem, fil):
        c = ctypes.c_void_p()
        assert L.kunCreateStreamEx(ex, modu, num_stocks, ctypes.byref(extra), ctypes.byref(c)) == KUN_SUCCESS
        resumed.append(c)
    bad = ExtraArgs(version=2, init_kind=KUN_INIT_NONE)
    c = ctypes.c_void_p()
    assert L.kunCreateStreamEx(ex, modu, num_stocks, ctypes.byref(bad), ctypes.byref(c)) == KUN_INVALID_ARGUMENT
    for r in rows[12:]:
        want = _push_run(L, ctx, handles, r, outs)
        for c in resumed:
            got = _push_run(L, c, handles, r, outs)
            for k in outs:
                np.testing.assert_array_equal(got[k], want[k], err_msg=k)
        np.testing.assert_array_equal(want["alpha101"], _alpha101(r))
    assert np.isfinite(want["corr10"]).all() and np.isfinite(want["ema5"]).all()
    for c in resumed:
        L.kunDestoryStream(c)
    L.kunDestoryStream(ctx)
    L.kunUnloadLibrary(lib)
    L.kunDestoryExecutor(ex)
