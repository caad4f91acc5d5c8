"""JIT driver: compiles factor Functions into a device module library with
hipcc and loads it (API of KunQuant/jit/cfake.py:127-211: compileit,
CppCompilerConfig)."""
import dataclasses
import os
import shutil
import tempfile
from dataclasses import dataclass, field
from typing import List, Tuple, Union

from .. import build as _build
from ..Driver import KunCompilerConfig
from ..Driver import compileit as driver_compileit
from ..Stage import Function
from ..codegen.hip import emit_module
from ..passes import Util
from ..runner import KunRunner


@dataclass
class CppCompilerConfig:
    """hipcc options. `machine` and `fast_linker_threads` exist for API
    compatibility and are ignored; the target is always gfx950."""
    opt_level: int = 3
    machine: object = None
    for_each: object = None
    other_flags: Tuple[str] = ()
    fast_linker_threads: int = 0
    compiler: str = "hipcc"
    obj_ext: str = "o"
    dll_ext: str = "so"
    jobs: int = 8


class NativeCPUFlags:
    pass


@dataclass
class X64CPUFlags:
    avx2: bool = True
    fma: bool = True
    avx512: bool = False
    avx512dq: bool = False
    avx512vl: bool = False


def get_runtime_path() -> str:
    return os.path.dirname(KunRunner.getRuntimePath())


def is_windows() -> bool:
    return False


def generate_sources(func: List[Tuple[str, Function, KunCompilerConfig]]) -> List[str]:
    """HIP translation units, one per module."""
    return [emit_module(driver_compileit(f, name, **dataclasses.asdict(cfg))) for name, f, cfg in func]


def compile_library(func, out_path: str, compiler_config: CppCompilerConfig = None, workdir: str = None) -> str:
    cfg = compiler_config or CppCompilerConfig()
    srcs = generate_sources(func)
    tmp = workdir or tempfile.mkdtemp(prefix="kunhip_")
    try:
        return _build.compile_module_sources(srcs, out_path, tmp, jobs=cfg.jobs, extra_flags=list(cfg.other_flags))
    finally:
        if not workdir and not Util.jit_debug_mode:
            shutil.rmtree(tmp, ignore_errors=True)


def compileit(func: List[Tuple[str, Function, KunCompilerConfig]], libname: str,
              compiler_config: CppCompilerConfig = None, tempdir: str = None, keep_files: bool = False,
              load: bool = True) -> Union["KunRunner.Library", str]:
    if keep_files and not tempdir:
        raise RuntimeError("if keep_files=True, tempdir should not be empty")
    base = tempdir or tempfile.mkdtemp(prefix="kunhip_")
    outdir = os.path.join(base, libname) if keep_files else base
    os.makedirs(outdir, exist_ok=True)
    out = os.path.join(outdir, f"{libname}.so")
    compile_library(func, out, compiler_config, workdir=os.path.join(outdir, "src"))
    if not load:
        return out
    return KunRunner.Library.load(out)
