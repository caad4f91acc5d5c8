"""Stream-state wire format on the device against the reference's own blobs
(tests/golden/gen_ref_vectors.py).

  * a context fed the same rows serialises to the reference's blob: same
    length, and every part equal (rows and state fields bit-exact with
    NaN payloads ignored, positions, counters, skip-list ranks and pairs
    exact);
  * a context resumed from the reference's blob continues exactly like the
    reference's uninterrupted run (its outputs), and serialises to the
    reference's later blob.

Bar: bit-exact (NaN payload bits ignored)."""
import os

import numpy as np
import pytest

from kunquant_amd.codegen.wire import WS_SKIPLIST, module_wire_table, parse_state
from kunquant_amd.runner import KunRunner as kr
from kunquant_amd.workloads import catalog

pytestmark = pytest.mark.gpu

GOLD = os.path.join(os.path.dirname(__file__), "golden")


@pytest.fixture(scope="module")
def executor():
    return kr.createSingleThreadExecutor()


def _same(a, b):
    return np.array_equal(a, b, equal_nan=True)


def _feed(st, rows, t0, t1, outs):
    """rows: {input: [T, S]}; runs ticks t0..t1-1, returns {out: [t1-t0, S]}."""
    hin = {k: st.queryBufferHandle(k) for k in rows}
    hout = {k: st.queryBufferHandle(k) for k in outs}
    res = {k: [] for k in outs}
    for t in range(t0, t1):
        for k, v in rows.items():
            st.pushData(hin[k], np.ascontiguousarray(v[t]))
        st.run()
        for k in outs:
            res[k].append(st.getCurrentBuffer(hout[k]).copy())
    return {k: np.array(v) for k, v in res.items()}


def _assert_same_state(m, S, ours, ref, table):
    assert len(ours) == len(ref)
    ob, os_, _ = parse_state(m, ours, S, table)
    rb, rs, _ = parse_state(m, ref, S, table)
    for x, y in zip(ob, rb):
        assert _same(x["rows"], y["rows"]), f"buffer {x['name']}"
        assert (x["pos"] == y["pos"]).all(), f"positions of {x['name']}"
    for j, (x, y) in enumerate(zip(os_, rs)):
        for f in x:
            if f in ("values", "index"):
                for s in range(S):
                    assert _same(x[f][s], y[f][s]), f"state {j} {f} stock {s}"
            elif f not in ("kind", "window"):
                assert _same(x[f], y[f]), f"state {j} field {f}"
        if x["kind"] == WS_SKIPLIST:
            assert (x["last"] == y["last"]).all()


@pytest.mark.parametrize("name", ["stream_f32", "stream_f64"])
def test_stream_graph_blobs(executor, name):
    g = np.load(os.path.join(GOLD, f"ref_{name}.npz"))
    _, mod = catalog.load(name)
    m = catalog.module_ir(name)
    table = module_wire_table(m)
    S, T = (int(x) for x in g["meta"])
    rows = {"a": g["in_a"].astype(mod.dtype)}
    outs = mod.getOutputNames()
    st = kr.StreamContext(executor, mod, S)
    _feed(st, rows, 0, 37, outs)
    _assert_same_state(m, S, st.serializeStates(), g["state37"].tobytes(), table)
    # resume from the reference's blob
    st2 = kr.StreamContext(executor, mod, S, g["state37"].tobytes())
    got = _feed(st2, rows, 37, T, outs)
    for k in outs:
        assert _same(got[k], g[f"out_{k}"][37:]), k
    _assert_same_state(m, S, st2.serializeStates(), g["state60"].tobytes(), table)


def test_serving_graph_blobs(executor):
    g = np.load(os.path.join(GOLD, "ref_wire_serve_f32.npz"))
    _, mod = catalog.load("serve_f32")
    m = catalog.module_ir("serve_f32")
    table = module_wire_table(m)
    S, T, K1 = (int(x) for x in g["meta"])
    rows = {k: g[f"in_{k}"] for k in mod.getInputNames()}
    outs = mod.getOutputNames()
    st = kr.StreamContext(executor, mod, S)
    _feed(st, rows, 0, K1, outs)
    _assert_same_state(m, S, st.serializeStates(), g["state_k1"].tobytes(), table)
    st2 = kr.StreamContext(executor, mod, S, g["state_k1"].tobytes())
    got = _feed(st2, rows, K1, T, outs)
    for k in outs:
        assert _same(got[k][-3:], g[f"out_{k}"]), k
    _assert_same_state(m, S, st2.serializeStates(), g["state_end"].tobytes(), table)


@pytest.mark.parametrize("name", ["stream_f32", "stream_f64"])
def test_pool_session_blobs(executor, name):
    """A pool session resumed from the reference's blob continues like the
    reference's run and serialises to its later blob; a session fed the
    rows from the start serialises to the reference's blob (next to a
    second, unrelated session of the same pool)."""
    g = np.load(os.path.join(GOLD, f"ref_{name}.npz"))
    _, mod = catalog.load(name)
    m = catalog.module_ir(name)
    table = module_wire_table(m)
    S, T = (int(x) for x in g["meta"])
    a = g["in_a"].astype(mod.dtype)
    outs = mod.getOutputNames()
    pool = kr.StreamPool(executor, mod, capacity_stocks=256, max_sessions=4)
    other = pool.admit(2 * S)
    fresh = pool.admit(S)
    resumed = pool.admit(S, state=g["state37"].tobytes())
    noise = np.random.default_rng(5).random((T, 2 * S)).astype(mod.dtype)
    for t in range(T):
        batch = [(fresh, {"a": a[t]}), (other, {"a": noise[t]})]
        if t >= 37:
            batch.append((resumed, {"a": a[t]}))
        pool.done(pool.submit(batch), wait=True)
        if t >= 37:
            for k in outs:
                assert _same(pool.read(resumed, k), g[f"out_{k}"][t]), (k, t)
        if t + 1 == 37:
            _assert_same_state(m, S, pool.serializeStates(fresh), g["state37"].tobytes(), table)
    _assert_same_state(m, S, pool.serializeStates(fresh), g["state60"].tobytes(), table)
    _assert_same_state(m, S, pool.serializeStates(resumed), g["state60"].tobytes(), table)
    with pytest.raises(RuntimeError, match="Failed to deserialize state buffer"):
        pool.admit(S, state=g["state37"].tobytes()[:-10])
    assert pool.admit(S) is not None   # the failed admission returned its lanes
