"""HIP emitter and driver checks that need no GPU."""
import pytest

from kunquant_amd.Driver import KunCompilerConfig, compileit
from kunquant_amd.Op import Builder, Input, Output, Rank
from kunquant_amd.Stage import Function
from kunquant_amd.codegen.hip import LDS_LIMIT, StageEmitter, emit_module
from kunquant_amd.ops import DiffWithWeightedSum, ExpMovingAvg, WindowedSum
from kunquant_amd.workloads import catalog


@pytest.mark.parametrize("name", sorted(catalog.CATALOG))
def test_emit_catalog(name):
    m = catalog.module_ir(name)
    src = emit_module(m)
    assert f"kun_module_{name}" in src
    for st in m.stages:
        if st.kind == "ts":
            em = StageEmitter(m, st, "k")
            assert em.lds_bytes <= LDS_LIMIT


def test_driver_option_errors():
    def g():
        b = Builder()
        with b:
            Output(WindowedSum(Input("a"), 5), "o")
        return Function(b.ops)
    with pytest.raises(RuntimeError, match="opt_reduce in stream mode"):
        compileit(g(), "m", input_layout="STREAM", output_layout="STREAM", options={"opt_reduce": True})
    with pytest.raises(RuntimeError, match="no_fast_stat=False is not supported in stream mode"):
        compileit(g(), "m", output_layout="STREAM", options={"no_fast_stat": False})
    with pytest.raises(RuntimeError, match="Bad dtype"):
        compileit(g(), "m", dtype="half")
    with pytest.raises(RuntimeError, match="allow_unaligned in stream mode"):
        compileit(g(), "m", output_layout="STREAM", allow_unaligned=True)


def test_generic_cross_sectional_rejected():
    class Diff(DiffWithWeightedSum):
        pass
    b = Builder()
    with b:
        Output(Diff(Input("a"), Input("b")), "o")
    with pytest.raises(RuntimeError, match="cannot run on the MI355X backend"):
        compileit(Function(b.ops), "m", input_layout="TS", output_layout="TS")


def test_ema_init_rejected_in_stream():
    b = Builder()
    with b:
        init = Input("__init_x")
        init.attrs["single_value"] = True
        Output(ExpMovingAvg(Input("a"), 5, init), "o")
    m = compileit(Function(b.ops), "m", input_layout="STREAM", output_layout="STREAM")
    with pytest.raises(RuntimeError, match="EMA with init_val is not supported in stream mode"):
        emit_module(m)
