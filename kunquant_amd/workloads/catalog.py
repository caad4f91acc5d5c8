"""Named modules that are compiled ahead of time by __graft_entry__.build()
(device code is never compiled inside a test or a GPU run).

Each entry: name -> (graph factory, KunCompilerConfig kwargs)."""
import os
from typing import Callable, Dict, Tuple

from ..Driver import KunCompilerConfig
from ..Op import Builder, Input, Output, Rank, Scale
from ..Stage import Function
from ..ops import (BackRef, DecayLinear, ExpMovingAvg, TsArgMax, TsArgMin, TsRank, WindowedAvg,
                   WindowedCorrelation, WindowedCovariance, WindowedKurt, WindowedLinearRegressionRSqaure,
                   WindowedLinearRegressionResi, WindowedLinearRegressionSlope, WindowedMax,
                   WindowedMaxDrawdown, WindowedMin, WindowedQuantile, WindowedSkew, WindowedStddev, WindowedSum,
                   Log, Exp, Sqrt, Sign, Abs, SetInfOrNanToValue, Pow, Clip)
from ..ops.ElewiseOp import GreaterThan, LessEqual, Select, Max, Min, Not, And, Or, Equals
from ..Op import ConstantOp
from .. import PACKAGE_DIR
from .graphs import random_graph, simple_graph

MODULE_DIR = os.path.join(PACKAGE_DIR, "lib", "modules")


def ops_graph() -> Function:
    """One output per primitive family, over inputs a and b."""
    b = Builder()
    with b:
        a, c = Input("a"), Input("b")
        Output(WindowedSum(a, 10), "sum10")
        Output(WindowedAvg(a, 7), "avg7")
        Output(WindowedStddev(a, 20), "std20")
        Output(WindowedCovariance(a, 10, c), "cov10")
        Output(WindowedCorrelation(a, 10, c), "corr10")
        Output(WindowedMin(a, 5), "min5")
        Output(WindowedMax(a, 5), "max5")
        Output(TsArgMax(a, 5), "argmax5")
        Output(TsArgMin(a, 6), "argmin6")
        Output(TsRank(a, 7), "tsrank7")
        Output(DecayLinear(a, 9), "decay9")
        Output(ExpMovingAvg(a, 5), "ema5")
        Output(WindowedQuantile(a, 10, 0.49), "q49")
        Output(WindowedQuantile(a, 12, 0.8), "q80")
        Output(WindowedLinearRegressionSlope(a, 10), "slope10")
        Output(WindowedLinearRegressionRSqaure(a, 10), "rsq10")
        Output(WindowedLinearRegressionResi(a, 10), "resi10")
        Output(WindowedSkew(a, 5), "skew5")
        Output(WindowedKurt(a, 6), "kurt6")
        Output(WindowedMaxDrawdown(a, 5), "mdd5")
        Output(BackRef(a, 3), "delay3")
        Output(a - BackRef(a, 2), "delta2")
        Output(Rank(a), "rank")
        Output(Scale(c), "scale")
        Output(Log(a), "log")
        Output(Exp(c), "exp")
        Output(Sqrt(a), "sqrt")
        Output(Sign(a - c), "sign")
        Output(SetInfOrNanToValue(a / c, 7.0), "finite")
        Output(Select(GreaterThan(a, c), a, c * 2), "select")
        Output(Max(a, c) - Min(a, c), "maxmin")
        Output(Pow(a, ConstantOp(3)), "pow3")
        Output(Pow(a, ConstantOp(0.5)), "pow05")
        Output(Clip(a - c, 0.3), "clip")
        Output(Select(And(Not(LessEqual(a, c)), Or(Equals(a, a), Equals(c, c))), a, c), "logic")
    return Function(b.ops)


def stream_graph() -> Function:
    """Stream-mode checks: quantile, EMA, regression slope, window stats and
    a cross-sectional rank."""
    b = Builder()
    with b:
        a = Input("a")
        Output(WindowedQuantile(a, 10, 0.49), "quantile")
        Output(ExpMovingAvg(a, 10), "ema")
        Output(WindowedLinearRegressionSlope(a, 10), "slope")
        Output(WindowedStddev(a, 10), "std")
        Output(TsRank(a, 5), "tsrank")
        Output(Rank(WindowedSum(a, 5)), "rank")
    return Function(b.ops)


CATALOG: Dict[str, Tuple[Callable[[], Function], dict]] = {
    "simple_f32": (simple_graph, dict(input_layout="TS", output_layout="TS")),
    "ops_f32": (ops_graph, dict(input_layout="TS", output_layout="TS")),
    "ops_f64": (ops_graph, dict(dtype="double", input_layout="TS", output_layout="TS",
                                options={"no_fast_stat": "no_warn"})),
    "ops_f32_sts": (ops_graph, dict(input_layout="STs", output_layout="STs")),
    "cfg2_f32": (lambda: random_graph(100, 102), dict(input_layout="TS", output_layout="TS")),
    "cfg2_f64": (lambda: random_graph(100, 102), dict(dtype="double", input_layout="TS", output_layout="TS",
                                                        options={"no_fast_stat": "no_warn"})),
    "stream_f64": (stream_graph, dict(dtype="double", input_layout="STREAM", output_layout="STREAM")),
    "stream_f32": (stream_graph, dict(input_layout="STREAM", output_layout="STREAM",
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
    from ..build import compile_module_sources
    from ..codegen.hip import emit_module
    os.makedirs(MODULE_DIR, exist_ok=True)
    for name in (names or CATALOG):
        src = emit_module(module_ir(name))
        compile_module_sources([src], library_path(name), os.path.join(MODULE_DIR, "_src", name), jobs=jobs)
