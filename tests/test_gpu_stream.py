"""Stream mode on the device: reference golden vectors, oracle parity over
NaN ticks, stream == batch, and state (de)serialisation (test_stream.py)."""
import os
import tempfile

import numpy as np
import pytest

from kunquant_amd.codegen.program import program_text
from kunquant_amd.runner import KunRunner as kr
from kunquant_amd.workloads import catalog
from oracle.kun_oracle import OracleProgram, OracleStream

pytestmark = pytest.mark.gpu
GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


@pytest.fixture(scope="module")
def executor():
    return kr.createSingleThreadExecutor()


def feed(st, rows, outs):
    ha = st.queryBufferHandle("a")
    hs = {k: st.queryBufferHandle(k) for k in outs}
    res = {k: np.empty_like(rows) for k in outs}
    for i, r in enumerate(rows):
        st.pushData(ha, np.ascontiguousarray(r))
        st.run()
        for k in outs:
            res[k][i] = st.getCurrentBuffer(hs[k])
    return res


@pytest.mark.parametrize("name", ["stream_f64", "stream_f32"])
def test_stream_matches_reference_golden(executor, name):
    z = np.load(os.path.join(GOLDEN, f"ref_{name}.npz"))
    S = int(z["meta"][0])
    _, mod = catalog.load(name)
    outs = mod.getOutputNames()
    got = feed(kr.StreamContext(executor, mod, S), z["in_a"], outs)
    for k in outs:
        assert np.array_equal(got[k], z[f"out_{k}"], equal_nan=True), k


@pytest.mark.parametrize("name", ["stream_f64", "stream_f32"])
def test_stream_matches_oracle_with_nans(executor, name):
    _, mod = catalog.load(name)
    S = 40
    rows = np.random.default_rng(31).random((90, S)).astype(mod.dtype)
    rows[np.random.default_rng(32).random(rows.shape) < 0.05] = np.nan
    outs = mod.getOutputNames()
    got = feed(kr.StreamContext(executor, mod, S), rows, outs)
    o = OracleStream(OracleProgram(program_text(catalog.module_ir(name))), S)
    for i, r in enumerate(rows):
        o.push("a", r)
        o.run()
        for k in outs:
            assert np.array_equal(got[k][i], o.get(k), equal_nan=True), (k, i)


def test_stream_equals_batch(executor):
    _, smod = catalog.load("stream_f64")
    _, bmod = catalog.load("streamgraph_batch_f64")
    S, T = 24, 80
    rows = np.random.default_rng(33).random((T, S))
    got = feed(kr.StreamContext(executor, smod, S), rows, smod.getOutputNames())
    batch = kr.runGraph(executor, bmod, {"a": rows}, 0, T)
    for k in got:
        assert np.array_equal(got[k], batch[k], equal_nan=True), k


def test_serialize_resume(executor):  # test_stream.py:53-126
    _, mod = catalog.load("stream_f64")
    S = 24
    a = np.random.default_rng(34).random((100, S))
    outs = mod.getOutputNames()
    ref = feed(kr.StreamContext(executor, mod, S), a, outs)
    st = kr.StreamContext(executor, mod, S)
    feed(st, a[:50], outs)
    path = os.path.join(tempfile.gettempdir(), "kunhip_stream_state.bin")
    st.serializeStates(path)
    blob = st.serializeStates()
    with open(path, "rb") as f:
        assert f.read() == blob
    del st
    for init in (blob, path):
        st2 = kr.StreamContext(executor, mod, S, init)
        tail = feed(st2, a[50:], outs)
        for k in outs:
            assert np.array_equal(tail[k], ref[k][50:], equal_nan=True), k
    with pytest.raises(RuntimeError) as e:
        kr.StreamContext(executor, mod, S, blob[:-10])
    assert str(e.value) == "Failed to deserialize state buffer"
    with open(path, "wb") as f:
        f.write(blob[:20])
    with pytest.raises(RuntimeError) as e:
        kr.StreamContext(executor, mod, S, path)
    assert str(e.value) == "Failed to read initial stream buffer"
    os.remove(path)
