"""Build helpers: the device runtime (libkunhip.so) and device module
libraries, compiled in-tree with hipcc for gfx950."""
import os
import shutil
import subprocess
from concurrent.futures import ThreadPoolExecutor
from typing import List, Optional, Sequence

from . import PACKAGE_DIR

CSRC = os.path.join(PACKAGE_DIR, "csrc")
LIB_DIR = os.path.join(PACKAGE_DIR, "lib")
RUNTIME_LIB = os.path.join(LIB_DIR, "libkunhip.so")
MODULE_DIR = os.path.join(LIB_DIR, "modules")
ARCH = os.environ.get("KUN_OFFLOAD_ARCH", "gfx950")

RUNTIME_SOURCES = ["runtime.hip", "stream.hip", "stream_pool.hip", "cross_section.hip", "row_reduce.hip",
                   "noalloc.hip", "capi.cpp"]


def hipcc() -> str:
    path = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
    if not os.path.exists(path):
        raise RuntimeError("hipcc not found: the MI355X backend needs ROCm's hipcc to build device code")
    return path


def hip_flags(extra: Sequence[str] = ()) -> List[str]:
    """Strict-IEEE device flags: no contraction (only explicit fma), no
    fast-math, correctly rounded fp32 div/sqrt, denormals kept, packed fp32
    disabled."""
    return [f"--offload-arch={ARCH}", "-O3", "-fPIC", "-std=c++17", "-ffp-contract=off", "-fno-fast-math",
            "-fhip-fp32-correctly-rounded-divide-sqrt", "-fno-gpu-flush-denormals-to-zero",
            "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
            "-Wno-unused-command-line-argument", "-I", CSRC] + list(extra)


def _run(cmd: List[str], cwd: Optional[str] = None):
    proc = subprocess.run(cmd, cwd=cwd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    if proc.returncode != 0:
        raise RuntimeError("command failed: " + " ".join(cmd) + "\n" + proc.stdout[-6000:])
    return proc.stdout


def build_runtime(force: bool = True) -> str:
    os.makedirs(LIB_DIR, exist_ok=True)
    srcs = [os.path.join(CSRC, s) for s in RUNTIME_SOURCES]
    if not force and os.path.exists(RUNTIME_LIB):
        newest = max(os.path.getmtime(p) for p in srcs + [os.path.join(CSRC, h) for h in os.listdir(CSRC)])
        if os.path.getmtime(RUNTIME_LIB) >= newest:
            return RUNTIME_LIB
    objs = []
    tmpdir = os.path.join(LIB_DIR, "_obj")
    os.makedirs(tmpdir, exist_ok=True)

    def compile_one(src):
        obj = os.path.join(tmpdir, os.path.basename(src) + ".o")
        lang = [] if src.endswith(".hip") else ["-x", "hip"]
        _run([hipcc()] + hip_flags() + lang + ["-c", src, "-o", obj])
        return obj

    with ThreadPoolExecutor(max_workers=4) as pool:
        objs = list(pool.map(compile_one, srcs))
    tmp = RUNTIME_LIB + ".tmp"
    _run([hipcc(), f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp] + objs + ["-ldl"])
    os.replace(tmp, RUNTIME_LIB)
    return RUNTIME_LIB


def compile_unit(src: str, workdir: str, i: int, extra_flags: Sequence[str] = ()) -> str:
    """Write generated translation unit `i` of a module and compile it."""
    os.makedirs(workdir, exist_ok=True)
    path = os.path.join(workdir, f"part{i}.hip")
    with open(path, "w") as f:
        f.write(src)
    obj = path + ".o"
    _run([hipcc()] + hip_flags(extra_flags) + ["-c", path, "-o", obj])
    return obj


def link_module(objs: List[str], out_path: str) -> str:
    tmp = out_path + ".tmp"
    _run([hipcc(), f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp] + list(objs))
    os.replace(tmp, out_path)
    return out_path


def compile_module_sources(sources: List[str], out_path: str, workdir: str, jobs: int = 8,
                           extra_flags: Sequence[str] = ()) -> str:
    """Compile generated module translation units and link one library."""
    os.makedirs(workdir, exist_ok=True)

    def compile_one(i_src):
        i, src = i_src
        path = os.path.join(workdir, f"part{i}.hip")
        with open(path, "w") as f:
            f.write(src)
        obj = path + ".o"
        _run([hipcc()] + hip_flags(extra_flags) + ["-c", path, "-o", obj])
        return obj

    with ThreadPoolExecutor(max_workers=max(1, jobs)) as pool:
        objs = list(pool.map(compile_one, enumerate(sources)))
    tmp = out_path + ".tmp"
    _run([hipcc(), f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp] + objs)
    os.replace(tmp, out_path)
    return out_path
