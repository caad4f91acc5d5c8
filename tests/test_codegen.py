"""HIP emitter and driver checks that need no GPU."""
import os

import pytest

from kunquant_amd.Driver import KunCompilerConfig, compileit
from kunquant_amd.Op import Builder, Input, Output, Rank
from kunquant_amd.Stage import Function
from kunquant_amd.codegen.hip import LDS_LIMIT, MAX_COMPS, LevelKernel, StageEmitter, emit_module, stage_plan
from kunquant_amd.codegen.plan import PAR, REC_OPS, SEQ, plan_stage
from kunquant_amd.ops import DiffWithWeightedSum, ExpMovingAvg, WindowedSum
from kunquant_amd.workloads import catalog


@pytest.mark.parametrize("name", sorted(catalog.CATALOG))
def test_emit_catalog(name):
    m = catalog.module_ir(name)
    src = emit_module(m)
    assert f"kun_module_{name}" in src
    for st in m.stages:
        if st.kind != "ts":
            continue
        if m.stream_mode:
            StageEmitter(m, st, "k")
            continue
        plan = stage_plan(m, st)
        for j, level in enumerate(plan.levels):
            # the emitter launches a level in groups of at most MAX_COMPS
            # components (levels of large libraries have more)
            for g in range(0, len(level.comps), MAX_COMPS):
                group = level.comps[g:g + MAX_COMPS]
                assert LevelKernel(m, st, plan, level, f"k{j}_{g}", None, group).lds_bytes <= LDS_LIMIT


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


def test_generic_cross_sectional_body_compiles_to_device_code(tmp_path):
    """A user's GenericCrossSectionalOp: its head and body become a device
    kernel (one thread per row) that hipcc compiles for gfx950; the oracle,
    which cannot run a user body, refuses the program."""
    from kunquant_amd.build import compile_unit
    from kunquant_amd.codegen.hip import emit_module_parts
    from kunquant_amd.codegen.program import program_text
    from kunquant_amd.ops import GenericCrossSectionalOp
    from oracle.kun_oracle import OracleProgram

    class Demean(GenericCrossSectionalOp):
        def __init__(self, v):
            super().__init__([v], None)

        def generate_head(self):
            return "T mean = 0;"

        def generate_body(self):
            return """
            for (size_t i = 0; i < num_stocks; ++i) mean = mean + input_0[i];
            mean = mean / (T)num_stocks;
            for (size_t i = 0; i < num_stocks; ++i) output_0[i] = input_0[i] - mean;
            """
    b = Builder()
    with b:
        Output(Demean(Input("a")), "o")
    m = compileit(Function(b.ops), "m", input_layout="TS", output_layout="TS")
    assert [s.kind for s in m.stages] == ["generic"]
    src = emit_module_parts(m, 1)[0]
    assert "void kun_m_s0_cs(kun::CsArgs A)" in src and "mean = mean / (T)num_stocks;" in src
    assert os.path.exists(compile_unit(src, str(tmp_path), 0))
    with pytest.raises(RuntimeError, match="cannot run"):
        OracleProgram(program_text(m))


def test_ema_init_rejected_in_stream():
    b = Builder()
    with b:
        init = Input("__init_x")
        init.attrs["single_value"] = True
        Output(ExpMovingAvg(Input("a"), 5, init), "o")
    m = compileit(Function(b.ops), "m", input_layout="STREAM", output_layout="STREAM")
    with pytest.raises(RuntimeError, match="EMA with init_val is not supported in stream mode"):
        emit_module(m)


def _batch_stage_plans():
    for name in sorted(catalog.CATALOG):
        m = catalog.module_ir(name)
        if m.stream_mode:
            continue
        for st in m.stages:
            if st.kind == "ts":
                yield name, st, plan_stage(st.func.ops)


def test_plan_invariants():
    """Every op lands in one component; recurrences only on sequential
    levels; values crossing levels are materialised by a lower level."""
    from kunquant_amd.Op import ConstantOp, ForeachBackWindow, Input
    from kunquant_amd.codegen.plan import _replicable
    for name, st, plan in _batch_stage_plans():
        ops = st.func.ops
        # a single-level stage replicates row-local values of its inputs
        # into every group reading them
        rep = _replicable(ops) if plan.fallback is not None else set()
        # Outputs of row-local values read back by several groups
        rep |= {o for lv in plan.levels for c in lv.comps for o in c.shadow}
        for lv in plan.levels:
            for c in lv.comps:
                rep |= {o for o in c.ops if any(o in cc.shadow for cc in lv.comps)}
        seen = {}
        for li, level in enumerate(plan.levels):
            for ci, comp in enumerate(level.comps):
                members = set(comp.ops)
                for op in comp.ops:
                    if isinstance(op, (Input, ConstantOp)):
                        continue
                    assert op not in seen or op in rep, (name, st.idx)
                    seen[op] = li
                    cls = op.__class__.__name__
                    if cls in REC_OPS:
                        assert level.kind == SEQ
                    if isinstance(op, ForeachBackWindow) and plan.fallback is None:
                        assert level.kind == PAR
                for op in comp.ops:
                    for inp in op.inputs:
                        if inp in members or isinstance(inp, (Input, ConstantOp)):
                            continue
                        assert inp in comp.ext and inp in plan.mat, (name, st.idx, op, inp)
                        assert seen[inp] < li
        computed = [op for op in ops if not isinstance(op, (Input, ConstantOp))]
        assert set(seen) == set(computed)
        if plan.fallback is not None:
            assert plan.fallback == "accumulator", (name, st.idx, plan.fallback)


def test_plan_halo():
    """Halo of a PAR component = longest chain of in-kernel window offsets:
    Delay(WindowedMax(WindowedMin(x, 5), 7), 3) recomputes 6 + 3 rows; the
    window over the input itself is rebuilt from HBM and costs none."""
    from kunquant_amd.ops import BackRef, WindowedMax, WindowedMin
    b = Builder()
    with b:
        Output(BackRef(WindowedMax(WindowedMin(Input("x"), 5), 7), 3), "o")
    m = compileit(Function(b.ops), "halo", input_layout="TS", output_layout="TS")
    (st,) = [s for s in m.stages if s.kind == "ts"]
    plan = stage_plan(m, st)
    assert [lv.kind for lv in plan.levels] == [PAR]
    assert LevelKernel(m, st, plan, plan.levels[0], "k").halo == 9


def test_plan_levels_of_running_sum():
    """ts_mean feeding a window loop: PAR (nothing) -> SEQ running sum ->
    PAR loop over the materialised mean."""
    from kunquant_amd.ops import WindowedAvg, WindowedMax
    b = Builder()
    with b:
        Output(WindowedMax(WindowedAvg(Input("x"), 10), 5), "o")
    m = compileit(Function(b.ops), "lv", input_layout="TS", output_layout="TS")
    (st,) = [s for s in m.stages if s.kind == "ts"]
    plan = stage_plan(m, st)
    assert [lv.kind for lv in plan.levels] == [SEQ, PAR]
    assert len(plan.mat) == 1


def test_fp64_finalize_emission():
    """Option fp64_finalize: window-loop sums and regression sums are double
    accumulators rounded to float once; without it every accumulator is
    float; stream modules reject it."""
    from kunquant_amd.ops import WindowedLinearRegressionSlope, WindowedStddev

    def graph():
        b = Builder()
        with b:
            a = Input("a")
            Output(WindowedStddev(a, 20), "std")
            Output(WindowedLinearRegressionSlope(a, 10), "slope")
        return Function(b.ops)

    src = emit_module(compileit(graph(), "fin", input_layout="TS", output_layout="TS",
                                options={"fp64_finalize": True}))
    assert "kun::LinReg<double>" in src and "double r" in src
    plain = emit_module(compileit(graph(), "nofin", input_layout="TS", output_layout="TS"))
    assert "kun::LinReg<float>" in plain and "double r" not in plain
    with pytest.raises(RuntimeError, match="fp64_finalize"):
        emit_module(compileit(graph(), "sfin", input_layout="STREAM", output_layout="STREAM",
                              options={"fp64_finalize": True, "opt_reduce": False}))


def test_stream_phase_plan():
    """Pool launch plan of the 100-factor stream module: every stage in
    exactly one phase; phases alternate time-series / cross-sectional; a
    time-series stage's producers run in an earlier phase or, if they are
    time-series, earlier in the same phase; a cross-sectional stage's
    producers run in earlier phases."""
    from kunquant_amd.codegen.hip import STAGE_TS, stream_phases
    ir = catalog.module_ir("serve_f32")
    phases = stream_phases(ir)
    where = {}
    for p, (kind, sts) in enumerate(phases):
        assert kind == p % 2 or all((st.kind == STAGE_TS) == (kind == 0) for st in sts)
        for j, st in enumerate(sts):
            assert st.idx not in where
            where[st.idx] = (p, j)
            assert (st.kind == STAGE_TS) == (kind == 0)
    assert sorted(where) == sorted(st.idx for st in ir.stages)
    writer = {b.idx: st for st in ir.stages for b in st.out_bufs}
    for st in ir.stages:
        p, j = where[st.idx]
        for b in st.in_bufs:
            w = writer.get(b.idx)
            if w is None or w is st:
                continue
            pw, jw = where[w.idx]
            if st.kind == STAGE_TS and w.kind == STAGE_TS:
                assert (pw, jw) < (p, j)
            else:
                assert pw < p
