"""kunRunGraphDeviceNoAlloc (the batch run with no device allocation, its
arena a fixed array of the library) against kunRunGraphDevice on the same
device buffers, through the C ABI: bit-exact, on two calls in a row, the
second with a smaller input (the arena then holds the first call's
contents); a run too large for the arena fails with a clear error and
leaves the outputs untouched.

Bar: every output equal bit for bit (NaN positions equal)."""
import ctypes
import re

import numpy as np
import pytest

from kunquant_amd.build import RUNTIME_LIB
from kunquant_amd.runner import KunRunner as kr
from kunquant_amd.workloads import catalog, data

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


def _arena_bytes():
    import os
    hdr = open(os.path.join(os.path.dirname(RUNTIME_LIB), "..", "..", "include", "kun_capi.h")).read()
    m = re.search(r"#define KUN_NOALLOC_SCRATCH_BYTES \((\d+)u << (\d+)\)", hdr)
    return int(m.group(1)) << int(m.group(2))


@pytest.fixture(scope="module")
def capi():
    L = ctypes.CDLL(RUNTIME_LIB)
    vp, sz = ctypes.c_void_p, ctypes.c_size_t
    f = L.kunRunGraphDeviceNoAlloc
    f.restype, f.argtypes = ctypes.c_int, [vp, vp, vp, sz, sz, sz, sz, vp]
    L.kunGetLastError.restype = ctypes.c_char_p
    L.kunCreateBufferNameMap.restype = vp
    L.kunSetBufferNameMap.argtypes = [vp, ctypes.c_char_p, vp]
    L.kunDestoryBufferNameMap.argtypes = [vp]
    return L


def _device_case(mod, S, T, seed, dtype):
    d = data.ohlcv(S, T, seed=seed, dtype=dtype, nan_frac=0.1)
    tdt = torch.float32 if dtype == np.float32 else torch.float64
    dev_in = {k: torch.from_numpy(d[k]).cuda() for k in mod.getInputNames()}
    outs = lambda: {k: torch.full((T, S), 3.0, dtype=tdt, device="cuda") for k in mod.getOutputNames()}
    return dev_in, outs(), outs()


def _no_alloc(L, ex, mod, ptrs, S, T):
    m = L.kunCreateBufferNameMap()
    try:
        for k, p in ptrs.items():
            L.kunSetBufferNameMap(m, k.encode(), p)
        return L.kunRunGraphDeviceNoAlloc(ex._h, mod._h, m, S, T, 0, T, None)
    finally:
        L.kunDestoryBufferNameMap(m)


@pytest.mark.parametrize("module,dtype", [("cfg2_f32", np.float32), ("prec_f32", np.float32),
                                          ("alpha158_f64", np.float64)])
def test_no_alloc_matches_device_run(capi, module, dtype):
    _, mod = catalog.load(module)
    ex = kr.createSingleThreadExecutor()
    for S, T, seed in ((300, 200, 31), (130, 70, 32)):   # the second call smaller than the first
        dev_in, via_pool, via_arena = _device_case(mod, S, T, seed, dtype)
        torch.cuda.synchronize()
        base = {k: v.data_ptr() for k, v in dev_in.items()}
        kr.runGraphDevice(ex, mod, dict(base, **{k: v.data_ptr() for k, v in via_pool.items()}), S, T, 0, T)
        ex.synchronize()
        rc = _no_alloc(capi, ex, mod, dict(base, **{k: v.data_ptr() for k, v in via_arena.items()}), S, T)
        assert rc == 0, capi.kunGetLastError()
        ex.synchronize()
        for k in via_pool:
            a, b = via_arena[k].cpu().numpy(), via_pool[k].cpu().numpy()
            assert np.array_equal(a, b, equal_nan=True), f"{module} {S}x{T}: output {k} differs"


def test_no_alloc_rejects_runs_beyond_the_arena(capi):
    _, mod = catalog.load("cfg2_f32")
    S, T = 4096, 512
    need, _ = mod.device_bytes(S, T)
    assert need > _arena_bytes()
    ex = kr.createSingleThreadExecutor()
    dev_in, _, out = _device_case(mod, S, T, 33, np.float32)
    torch.cuda.synchronize()
    rc = _no_alloc(capi, ex, mod, dict({k: v.data_ptr() for k, v in dev_in.items()},
                                       **{k: v.data_ptr() for k, v in out.items()}), S, T)
    assert rc != 0
    assert b"fixed arena" in capi.kunGetLastError()
    ex.synchronize()
    assert all(bool((v == 3.0).all()) for v in out.values())
