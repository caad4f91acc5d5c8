"""The C ABI library without a GPU: it loads, exports every entry point
declared in include/kun_capi.h, loads prebuilt module libraries and answers
introspection queries."""
import ctypes
import os
import re

import pytest

from kunquant_amd.build import RUNTIME_LIB
from kunquant_amd.runner import KunRunner as kr
from kunquant_amd.workloads import catalog

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    text = open(os.path.join(ROOT, "include", "kun_capi.h")).read()
    return sorted(set(re.findall(r"KUN_API\s+[^;(]*?\b(kun\w+)\s*\(", text)))


def test_header_declares_reference_entry_points():
    names = set(declared_symbols())
    reference = {"kunCreateSingleThreadExecutor", "kunCreateMultiThreadExecutor", "kunDestoryExecutor",
                 "kunLoadLibrary", "kunGetModuleFromLibrary", "kunUnloadLibrary", "kunCreateBufferNameMap",
                 "kunDestoryBufferNameMap", "kunSetBufferNameMap", "kunEraseBufferNameMap", "kunRunGraph",
                 "kunCreateStream", "kunCreateStreamEx", "kunStreamSerializeStates", "kunQueryBufferHandle",
                 "kunStreamGetCurrentBuffer", "kunStreamPushData", "kunStreamRun", "kunDestoryStream"}
    assert reference <= names


def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(RUNTIME_LIB)
    missing = [n for n in declared_symbols() if not hasattr(lib, n)]
    assert not missing, missing


def test_prebuilt_modules_introspect():
    for name in catalog.CATALOG:
        lib, mod = catalog.load(name)
        ir = catalog.module_ir(name)
        assert mod.getOutputNames() == ir.output_names()
        assert mod.getOutputUnreliableCount() == {b.name: b.unreliable for b in ir.buffers if b.kind == "OUTPUT"}
        assert mod.output_layout == ir.output_layout
        assert mod.blocking_len == ir.blocking_len
        assert mod.program().startswith("KUNPROG 1")


def test_library_errors():
    with pytest.raises(RuntimeError, match="Cannot load library"):
        kr.Library.load(os.path.join(ROOT, "no_such_library.so"))
    lib, _ = catalog.load("simple_f32")
    with pytest.raises(RuntimeError, match="Module name not found"):
        lib.getModule("testRuntimeModule2")


def test_device_bytes_reuse_slots():
    """The memory plan of a batch run (host-side, no GPU): module temporaries
    and values handed between a stage's kernels share fewer [rows][stocks]
    slots than there are temporaries; a module without temporaries still
    gets slots for its materialised values only."""
    _, mod = catalog.load("cfg2_f32")
    nbytes, slots = mod.device_bytes(8192, 1024)
    ntemp = sum(1 for _, kind, _, _ in mod.buffers if kind == 2)
    assert 0 < slots < ntemp
    assert nbytes >= slots * 8192 * 1024 * 4
    _, small = catalog.load("simple_f32")
    nb2, slots2 = small.device_bytes(100, 50)
    assert slots2 <= 4 and nb2 >= slots2 * 100 * 50 * 4
