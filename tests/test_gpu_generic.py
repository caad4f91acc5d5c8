"""Generic cross-sectional op (a GenericCrossSectionalOp's body compiled as
device code, one thread per row) between time-series stages:
DiffWithWeightedSum(a - b, c), its 5-row average and its rank.

  * TS fp32 module: equal to the reference's own outputs
    (tests/golden/ref_generic_f32.npz) bit for bit;
  * STs fp64 module: equal to the oracle bit for bit;
  * stream module: tick by tick equal to the reference's batch outputs;
  * session pools refuse the module (the body runs in stream contexts).

Bar: bit-exact (NaN positions equal, NaN payload bits ignored)."""
import os

import numpy as np
import pytest

from kunquant_amd.codegen.program import program_text
from kunquant_amd.runner import KunRunner as kr
from kunquant_amd.workloads import catalog
from oracle.kun_oracle import OracleProgram

pytestmark = pytest.mark.gpu

GOLD = os.path.join(os.path.dirname(__file__), "golden")


@pytest.fixture(scope="module")
def executor():
    return kr.createSingleThreadExecutor()


@pytest.fixture(scope="module")
def golden():
    z = np.load(os.path.join(GOLD, "ref_generic_f32.npz"))
    ins = {k[3:]: z[k] for k in z.files if k.startswith("in_")}
    outs = {k[4:]: z[k] for k in z.files if k.startswith("out_")}
    return ins, outs, [int(x) for x in z["meta"]]


def _same(a, b):
    return np.array_equal(a, b, equal_nan=True)


def test_ts_module_matches_reference(executor, golden):
    ins, outs, (S, T, start, L) = golden
    _, mod = catalog.load("generic_f32")
    got = kr.runGraph(executor, mod, ins, start, L)
    bad = [k for k in outs if not _same(got[k], outs[k])]
    assert not bad, bad


def _to_sts(x, b):
    T, S = x.shape
    return np.ascontiguousarray(x.reshape(T, S // b, b).transpose(1, 0, 2))


def test_sts_module_matches_oracle(executor):
    _, mod = catalog.load("generic_f64_sts")
    S, T = 48, 40
    rng = np.random.default_rng(61)
    ins = {k: _to_sts(rng.random((T, S)), 4) for k in ("a", "b", "c")}
    ins["c"][2, 7, 1] = np.nan
    got = kr.runGraph(executor, mod, ins, 0, T)
    exp = OracleProgram(program_text(catalog.module_ir("generic_f64_sts"))).run(ins, S, T, 0, T)
    bad = [k for k in exp if not _same(got[k], exp[k])]
    assert not bad, bad


def test_stream_module_matches_reference(executor, golden):
    ins, outs, (S, T, start, L) = golden
    _, mod = catalog.load("generic_stream_f32")
    st = kr.StreamContext(executor, mod, S)
    h = {k: st.queryBufferHandle(k) for k in list(ins) + list(outs)}
    for t in range(T):
        for k in ins:
            st.pushData(h[k], np.ascontiguousarray(ins[k][t]))
        st.run()
        for k in outs:
            assert _same(st.getCurrentBuffer(h[k]), outs[k][t]), (k, t)


def test_pool_refuses_generic_stage(executor):
    _, mod = catalog.load("generic_stream_f32")
    with pytest.raises(RuntimeError, match="generic cross-sectional"):
        kr.StreamPool(executor, mod, capacity_stocks=128, max_sessions=2)
