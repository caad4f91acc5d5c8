"""Named modules that are compiled ahead of time by __graft_entry__.build()
(device code is never compiled inside a test or a GPU run).

Each entry: name -> (graph factory, KunCompilerConfig kwargs)."""
import os
from typing import Callable, Dict, Tuple

from .. import PACKAGE_DIR
from .graphs import (aligned_graph, alpha101_graph, large_window_graph, capi_batch_graph, capi_stream_graph, alpha158_graph, cross_section_graph, generic_cs_graph, tsrank_chain_graph, ops_graph, precision_graph, random_graph, simple_graph,
                     stream_graph)

MODULE_DIR = os.path.join(PACKAGE_DIR, "lib", "modules")


CATALOG: Dict[str, Tuple[Callable, dict]] = {
    "simple_f32": (simple_graph, dict(input_layout="TS", output_layout="TS")),
    "ops_f32": (ops_graph, dict(input_layout="TS", output_layout="TS")),
    "ops_f64": (ops_graph, dict(dtype="double", input_layout="TS", output_layout="TS",
                                options={"no_fast_stat": "no_warn"})),
    "ops_f32_sts": (ops_graph, dict(input_layout="STs", output_layout="STs")),
    "cfg2_f32": (lambda: random_graph(100, 102), dict(input_layout="TS", output_layout="TS")),
    "cfg2_f64": (lambda: random_graph(100, 102), dict(dtype="double", input_layout="TS", output_layout="TS",
                                                        options={"no_fast_stat": "no_warn"})),
    "rand40_f64": (lambda: random_graph(40, 103), dict(dtype="double", input_layout="TS", output_layout="TS",
                                                        options={"no_fast_stat": "no_warn"})),
    "streamgraph_batch_f64": (stream_graph, dict(dtype="double", input_layout="TS", output_layout="TS",
                                                  options={"no_fast_stat": True})),
    # the modules of the reference's C API test (names included)
    "testRuntimeModule": (capi_batch_graph, dict(input_layout="STs", output_layout="STs")),
    "alpha_101_stream": (capi_stream_graph, dict(partition_factor=8, input_layout="STREAM", output_layout="STREAM",
                                                 options={"opt_reduce": False, "fast_log": True})),
    "aligned_f32": (aligned_graph, dict(input_layout="TS", output_layout="TS", allow_unaligned=False)),
    "cs_f32": (cross_section_graph, dict(input_layout="TS", output_layout="TS")),
    # the Alpha158 factor library as the reference's test compiles it (fast
    # statistics on, as there; "no_warn" only silences the notice)
    "alpha158_f64": (alpha158_graph, dict(dtype="double", blocking_len=4, partition_factor=4, input_layout="TS",
                                          output_layout="TS", options={"opt_reduce": True, "fast_log": True,
                                                                       "no_fast_stat": "no_warn"})),
    # the Alpha101 library compiled as the reference's test compiles it
    # (fp32: fast log, rolling sums; fp64: fast statistics on as well)
    "alpha101_f32": (alpha101_graph, dict(input_layout="TS", output_layout="TS",
                                          options={"opt_reduce": True, "fast_log": True})),
    "alpha101_f64": (alpha101_graph, dict(dtype="double", input_layout="TS", output_layout="TS",
                                          options={"opt_reduce": True, "fast_log": True, "no_fast_stat": "no_warn"})),
    # 200-row windows: skip-list TsRank / Min / Max / ArgMin / ArgMax (fp64)
    "skiplist_f64": (large_window_graph, dict(dtype="double", input_layout="TS", output_layout="TS",
                                              options={"no_fast_stat": "no_warn"})),
    # rank of a rank (the reference's copy-constructed rank state)
    "tsrank_chain_f32": (tsrank_chain_graph, dict(input_layout="TS", output_layout="TS")),
    # generic cross-sectional op (device-compiled body) in TS, STs and stream modules
    "generic_f32": (generic_cs_graph, dict(input_layout="TS", output_layout="TS")),
    "generic_f64_sts": (generic_cs_graph, dict(dtype="double", input_layout="STs", output_layout="STs",
                                               options={"no_fast_stat": "no_warn"})),
    "generic_stream_f32": (generic_cs_graph, dict(input_layout="STREAM", output_layout="STREAM")),
    # graphs of the batched-jobs workload (64 uneven jobs draw from this pool)
    "jobs_g0": (lambda: random_graph(20, 4000), dict(input_layout="TS", output_layout="TS")),
    "jobs_g1": (lambda: random_graph(30, 4001), dict(input_layout="TS", output_layout="TS")),
    "jobs_g2": (lambda: random_graph(45, 4002), dict(input_layout="TS", output_layout="TS")),
    "jobs_g3": (lambda: random_graph(60, 4003), dict(input_layout="TS", output_layout="TS")),
    "cs_f64": (cross_section_graph, dict(dtype="double", input_layout="TS", output_layout="TS",
                                        options={"no_fast_stat": "no_warn"})),
    "stream_f64": (stream_graph, dict(dtype="double", input_layout="STREAM", output_layout="STREAM")),
    "stream_f32": (stream_graph, dict(input_layout="STREAM", output_layout="STREAM",
                                      options={"opt_reduce": False})),
    # precision run: 150 factors; reference fp32 semantics, fp32 with fp64
    # finalisation, and the fp64 reference of the same algorithm
    "prec_f32": (lambda: precision_graph(150, 3), dict(input_layout="TS", output_layout="TS")),
    "prec_f32x": (lambda: precision_graph(150, 3), dict(input_layout="TS", output_layout="TS",
                                                        options={"fp64_finalize": True})),
    "prec_f64": (lambda: precision_graph(150, 3), dict(dtype="double", input_layout="TS", output_layout="TS",
                                                       options={"no_fast_stat": True})),
    # serving workload: the 100-factor random graph in stream mode (one row per tick)
    "serve_f32": (lambda: random_graph(100, 5000), dict(input_layout="STREAM", output_layout="STREAM",
                                                        options={"opt_reduce": False})),
}


def module_ir(name: str):
    from ..Driver import compileit
    factory, cfg = CATALOG[name]
    return compileit(factory(), name, **dict(cfg, options=dict(cfg.get("options", {}))))


def library_path(name: str) -> str:
    return os.path.join(MODULE_DIR, f"{name}.so")


def load(name: str):
    """(KunRunner.Library, Module) of a prebuilt catalog module."""
    from ..runner import KunRunner
    path = library_path(name)
    if not os.path.exists(path):
        raise RuntimeError(f"module {name} is not built ({path}); run __graft_entry__.build()")
    lib = KunRunner.Library.load(path)
    return lib, lib.getModule(name)


def build_all(jobs: int = 8, names=None):
    """Compile every catalog module: large modules are split into several
    translation units, and all units of all modules share `jobs` hipcc
    processes (largest first), then each module links."""
    from concurrent.futures import ThreadPoolExecutor
    from ..build import compile_unit, link_module
    from ..codegen.hip import emit_module_parts
    os.makedirs(MODULE_DIR, exist_ok=True)
    todo = list(names or CATALOG)
    srcs = {}
    for name in todo:
        ir = module_ir(name)
        whole = emit_module_parts(ir, 1)[0]
        nparts = max(1, min(jobs, len(whole) // (2 << 20)))
        srcs[name] = emit_module_parts(ir, nparts + 1) if nparts > 1 else [whole]
    units = [(name, i, src) for name in todo for i, src in enumerate(srcs[name])]
    workdir = lambda name: os.path.join(MODULE_DIR, "_src", name)
    with ThreadPoolExecutor(max_workers=max(1, jobs)) as pool:
        objs = list(pool.map(lambda u: (u[0], u[1], compile_unit(u[2], workdir(u[0]), u[1])),
                             sorted(units, key=lambda u: -len(u[2]))))
    by_mod = {}
    for name, i, obj in objs:
        by_mod.setdefault(name, {})[i] = obj
    for name in todo:
        link_module([by_mod[name][i] for i in sorted(by_mod[name])], library_path(name))
