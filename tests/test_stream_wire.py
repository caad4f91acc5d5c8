"""Stream-state wire table against blobs the reference's own StreamContext
serialised (tests/golden/gen_ref_vectors.py: Runtime.cpp:534-546 through
kun::MemoryOutputStream), on the CPU.

The compiler's wire table must split each blob exactly (every byte
consumed), and the parts must hold what the mathematics fixes for the
inputs that were pushed:
  * input buffers: row p = input of the last tick t <= K-1 with t % window
    == p (NaN before the first tick), positions = K % window;
  * outputs of window 1 = the reference's outputs of tick K-1;
  * the quantile skip list of window 10 over `a` = the non-NaN values of
    the last 10 inputs, sorted (as doubles), time index 0 (stream mode);
  * EMA state = the EMA output of tick K-1; regression / rolling-sum NaN
    counters = NaNs in the window; regression count = non-NaN values;
    rolling sums = the window's sum to 1e-12.
The 100-factor serving graph's blobs exercise windowed stage temporaries
(<stage>_<op index> buffers) and buffers read back by their writer."""
import os

import numpy as np
import pytest

from kunquant_amd.codegen.wire import (WS_EMA, WS_FASTSUM, WS_LINREG, WS_SKIPLIST, module_wire_table,
                                       parse_state)
from kunquant_amd.workloads import catalog

GOLD = os.path.join(os.path.dirname(__file__), "golden")


def _load(name):
    return np.load(os.path.join(GOLD, f"ref_{name}.npz"))


def _window_rows(x, K, W):
    """Reference StreamBuffer rows after K ticks of pushing x[t]."""
    rows = np.full((W,) + x.shape[1:], np.nan, dtype=x.dtype)
    for t in range(max(0, K - W), K):
        rows[t % W] = x[t]
    return rows


def _same(a, b):
    return np.array_equal(a, b, equal_nan=True)


@pytest.mark.parametrize("name,key,K", [("stream_f32", "state37", 37), ("stream_f32", "state60", 60),
                                        ("stream_f64", "state37", 37), ("stream_f64", "state60", 60)])
def test_stream_graph_blob(name, key, K):
    g = _load(name)
    m = catalog.module_ir(name)
    S = int(g["meta"][0])
    blob = g[key].tobytes()
    table = module_wire_table(m)
    bufs, states, used = parse_state(m, blob, S, table)
    assert used == len(blob)
    a = g["in_a"]
    by = {b["name"]: b for b in bufs}
    W = by["a"]["window"]
    assert W == 11
    assert _same(by["a"]["rows"], _window_rows(a, K, W))
    for b in bufs:
        assert (b["pos"] == (K % b["window"] if b["window"] > 1 else 0)).all(), b["name"]
    outs = [k[4:] for k in g.files if k.startswith("out_")]
    for k in outs:
        assert by[k]["window"] == 1
        assert _same(by[k]["rows"][0], g[f"out_{k}"][K - 1]), k
    win = a[K - 10:K].astype(np.float64)
    kinds = [s["kind"] for s in states]
    sk = states[kinds.index(WS_SKIPLIST)]
    for s in range(S):
        col = win[:, s]
        assert _same(sk["values"][s], np.sort(col[~np.isnan(col)])), s
        assert (sk["index"][s] == 0).all()
    ema = states[kinds.index(WS_EMA)]
    assert _same(ema["v"], g["out_ema"][K - 1])
    lr = states[kinds.index(WS_LINREG)]
    nn = np.isnan(a[K - 10:K]).sum(axis=0)
    assert (lr["nnan"] == nn).all() and (lr["i"] == 10 - nn).all()
    for st in states:
        if st["kind"] == WS_FASTSUM:
            w = st["window"]
            x = a[K - w:K]
            assert (st["nnan"] == np.isnan(x).sum(axis=0)).all()
            ok = ~np.isnan(x).any(axis=0)
            np.testing.assert_allclose(st["v"][ok], x[:, ok].sum(axis=0), rtol=1e-12)
    assert WS_FASTSUM in kinds or name == "stream_f32"


@pytest.mark.parametrize("key", ["state_k1", "state_end"])
def test_serving_graph_blob(key):
    g = _load("wire_serve_f32")
    S, T, K1 = (int(x) for x in g["meta"])
    K = K1 if key == "state_k1" else T
    m = catalog.module_ir("serve_f32")
    blob = g[key].tobytes()
    bufs, states, used = parse_state(m, blob, S)
    assert used == len(blob)
    assert not states          # window ops of this graph run as loops over ring rows
    temps = [b for b in bufs if b["name"] not in {x.name for x in m.buffers}]
    assert len(temps) >= 10 and max(b["window"] for b in temps) >= 30
    for b in bufs:
        assert (b["pos"] == (K % b["window"] if b["window"] > 1 else 0)).all(), b["name"]
    by = {b["name"]: b for b in bufs}
    for k in m.input_names():
        x = g[f"in_{k}"]
        w = by[k]["window"]
        assert _same(by[k]["rows"], _window_rows(x, K, w)), k
    if key == "state_end":
        for k in m.output_names():
            assert _same(by[k]["rows"][by[k]["pos"][0] - 1 if by[k]["window"] > 1 else 0], g[f"out_{k}"][-1]), k


def test_truncated_blob_is_rejected():
    g = _load("stream_f32")
    m = catalog.module_ir("stream_f32")
    with pytest.raises(ValueError):
        parse_state(m, g["state37"].tobytes()[:-3], int(g["meta"][0]))
