"""CPU emulation of generated batch stage kernels (test infrastructure).

The HIP source a module library is built from is compiled with g++ against
a stand-in for <hip/hip_runtime.h> (tests/emu/include) and its stage kernels
are run lane by lane (tests/emu/kun_emu.cpp). Tests compare the result of
every time-series stage with the oracle's value of the same buffers, so the
emitter, the kernel plan and the row chunking are checked without a GPU."""
import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
BUILD = os.path.join(HERE, "_build")


def build(module_ir, lds_budget=None) -> str:
    from kunquant_amd.codegen.hip import emit_module, ident
    os.makedirs(BUILD, exist_ok=True)
    name = ident(module_ir.name) + ("" if lds_budget is None else f"_lds{lds_budget}")
    src = os.path.join(BUILD, f"{name}.cpp")
    out = os.path.join(BUILD, f"{name}.so")
    text = emit_module(module_ir, lds_budget)
    if os.path.exists(out) and os.path.exists(src) and open(src).read() == text:
        return out
    with open(src, "w") as f:
        f.write(text)
    cmd = ["g++", "-O1", "-std=c++17", "-ffp-contract=off", "-fno-fast-math", "-fPIC", "-shared", "-w",
           "-I", os.path.join(HERE, "include"), "-I", os.path.join(ROOT, "kunquant_amd", "csrc"),
           src, os.path.join(HERE, "kun_emu.cpp"), "-o", out + ".tmp"]
    subprocess.run(cmd, check=True)
    os.replace(out + ".tmp", out)
    return out


class EmuModule:
    def __init__(self, module_ir, lds_budget=None):
        from kunquant_amd.codegen.hip import ident
        self.ir = module_ir
        self.lib = ctypes.CDLL(build(module_ir, lds_budget))
        self.desc = ctypes.c_void_p.in_dll(self.lib, f"kun_module_{ident(module_ir.name)}")
        self.desc_ptr = ctypes.addressof(self.desc)
        f = self.lib.kun_emu_run_stage
        f.restype = ctypes.c_int
        f.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p, ctypes.c_long, ctypes.c_long,
                      ctypes.c_long, ctypes.c_long, ctypes.c_long]

    def run_stage(self, stage: int, bufs: dict, S: int, T: int, start: int, L: int, chunk_rows: int):
        """bufs: every module buffer by name (numpy arrays in the module's
        layouts); the stage's outputs are written in place."""
        ptrs = (ctypes.c_void_p * len(self.ir.buffers))()
        for b in self.ir.buffers:
            arr = bufs[b.name]
            assert arr.flags["C_CONTIGUOUS"]
            ptrs[b.idx] = arr.ctypes.data
        if self.lib.kun_emu_run_stage(self.desc_ptr, stage, ptrs, S, T, start, L, chunk_rows):
            raise RuntimeError("emulated stage failed")
