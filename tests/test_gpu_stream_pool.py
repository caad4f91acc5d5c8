"""Session pools on the device: every session's rows, tick by tick, equal
those of a stream context of its own fed the same rows -- alone and mixed
with other sessions in cycles of varying membership, with two cycles in
flight, and across admission, release and re-admission into freed lanes.
The 100-factor serving module is also checked against the oracle.

Stream modules are compiled aligned, so every universe is a multiple of
the blocking length (8 lanes in fp32).

Bar: bit-exact (NaN positions equal, NaN payload bits ignored)."""
import numpy as np
import pytest

from kunquant_amd.codegen.program import program_text
from kunquant_amd.runner import KunRunner as kr
from kunquant_amd.workloads import catalog, data
from oracle.kun_oracle import OracleProgram, OracleStream

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def executor():
    return kr.createSingleThreadExecutor()


def _ticks(mod, S, T, seed):
    """[T] of {input name: row [S]} (OHLCV with NaN cells)."""
    d = data.ohlcv(S, T, seed=seed, dtype=mod.dtype, nan_frac=0.05)
    names = mod.getInputNames()
    if names == ["a"]:
        d = {"a": d["close"]}
    return [{k: np.ascontiguousarray(d[k][t]) for k in names} for t in range(T)]


def _alone(executor, mod, S, rows, outs):
    st = kr.StreamContext(executor, mod, S)
    hin = {k: st.queryBufferHandle(k) for k in rows[0]}
    hout = {k: st.queryBufferHandle(k) for k in outs}
    res = []
    for r in rows:
        for k, v in r.items():
            st.pushData(hin[k], v)
        st.run()
        res.append({k: st.getCurrentBuffer(hout[k]).copy() for k in outs})
    return res


def _same(a, b):
    return np.array_equal(a, b, equal_nan=True)


def test_serving_module_matches_oracle(executor):
    _, mod = catalog.load("serve_f32")
    S, T = 72, 24
    rows = _ticks(mod, S, T, seed=81)
    outs = mod.getOutputNames()
    got = _alone(executor, mod, S, rows, outs)
    o = OracleStream(OracleProgram(program_text(catalog.module_ir("serve_f32"))), S)
    for t, r in enumerate(rows):
        for k, v in r.items():
            o.push(k, v)
        o.run()
        bad = [k for k in outs if not _same(got[t][k], o.get(k))]
        assert not bad, f"tick {t}: {bad[:3]}"


@pytest.mark.parametrize("name,sizes,T", [("stream_f32", (72, 8, 136, 256), 40),
                                          ("serve_f32", (72, 296, 128), 30)])
def test_pool_sessions_equal_alone(executor, name, sizes, T):
    """Sessions join cycles at random (a session with no tick this cycle
    sits out), two cycles in flight, one session admitted late."""
    _, mod = catalog.load(name)
    outs = mod.getOutputNames()
    rng = np.random.default_rng(82)
    feeds = [_ticks(mod, S, T, seed=90 + j) for j, S in enumerate(sizes)]
    alone = [_alone(executor, mod, S, f, outs) for S, f in zip(sizes, feeds)]
    pool = kr.StreamPool(executor, mod, capacity_stocks=sum(sizes) + 64 * len(sizes), max_sessions=8,
                         max_inflight=2)
    sids = [pool.admit(S) for S in sizes[:-1]] + [None]
    done = [0] * len(sizes)
    got = [[] for _ in sizes]
    pending = []  # (cycle, [session index])
    step = 0
    while min(done) < T:
        if sids[-1] is None and step == 7:
            sids[-1] = pool.admit(sizes[-1])
        step += 1
        members = [j for j in range(len(sizes)) if sids[j] is not None and done[j] < T
                   and not any(j in m for _, m in pending) and rng.random() < 0.7]
        if members:
            cyc = pool.submit([(sids[j], feeds[j][done[j]]) for j in members])
            pending.append((cyc, members))
        if len(pending) == 2 or (not members and pending):
            cyc, m = pending.pop(0)
            pool.done(cyc, wait=True)
            for j in m:
                got[j].append({k: pool.read(sids[j], k) for k in outs})
                done[j] += 1
    for cyc, m in pending:
        pool.done(cyc, wait=True)
    for j in range(len(sizes)):
        assert pool.ticks(sids[j]) == T
        for t in range(T):
            bad = [k for k in outs if not _same(got[j][t][k], alone[j][t][k])]
            assert not bad, f"{name} session {j} (S={sizes[j]}) tick {t}: {bad[:3]}"


def test_pool_admission_by_footprint(executor):
    """A session that does not fit waits; after a release it is admitted
    into the freed lanes and starts from fresh state."""
    _, mod = catalog.load("stream_f32")
    outs = mod.getOutputNames()
    pool = kr.StreamPool(executor, mod, capacity_stocks=256, max_sessions=4)
    a = pool.admit(128)
    b = pool.admit(96)
    assert a is not None and b is not None
    assert pool.admit(72) is None            # 64-lane ranges: 128 + 128 taken
    rows_a = _ticks(mod, 128, 5, seed=95)
    for r in rows_a:
        pool.done(pool.submit([(a, r)]), wait=True)
    pool.release(a)
    assert pool.free_stocks() == 128
    c = pool.admit(120)
    assert c is not None
    rows_c = _ticks(mod, 120, 12, seed=96)
    rows_b = _ticks(mod, 96, 12, seed=97)
    alone_c = _alone(executor, mod, 120, rows_c, outs)
    alone_b = _alone(executor, mod, 96, rows_b, outs)
    for t in range(12):
        pool.done(pool.submit([(c, rows_c[t]), (b, rows_b[t])]), wait=True)
        for k in outs:
            assert _same(pool.read(c, k), alone_c[t][k]), (t, k)
            assert _same(pool.read(b, k), alone_b[t][k]), (t, k)
    with pytest.raises(RuntimeError, match="twice"):
        pool.submit([(b, rows_b[0]), (b, rows_b[0])])


def test_pool_readback_and_phase_times(executor):
    """With readback buffers set, a finished cycle's results are in pinned
    host memory: they equal a stream context's outputs tick by tick (two
    cycles in flight, sessions sitting out at random). Profiled cycles give
    one time per phase plus input staging and readback."""
    from kunquant_amd.codegen.hip import stream_phases
    _, mod = catalog.load("serve_f32")
    outs = mod.getOutputNames()
    sizes, T = (72, 128, 40), 10
    rng = np.random.default_rng(83)
    feeds = [_ticks(mod, S, T, seed=70 + j) for j, S in enumerate(sizes)]
    alone = [_alone(executor, mod, S, f, outs) for S, f in zip(sizes, feeds)]
    pool = kr.StreamPool(executor, mod, capacity_stocks=512, max_sessions=4, max_inflight=2)
    sids = [pool.admit(S) for S in sizes]
    pool.set_readback(outs)
    pool.profile(True)
    done, pending, times = [0] * 3, [], []
    while min(done) < T:
        members = [j for j in range(3) if done[j] < T and not any(j in m for _, m in pending)
                   and rng.random() < 0.8]
        if members:
            pending.append((pool.submit([(sids[j], feeds[j][done[j]]) for j in members]), members))
        if len(pending) == 2 or (not members and pending):
            cyc, m = pending.pop(0)
            assert pool.done(cyc, wait=True)
            for j in m:
                bad = [k for k in outs if not _same(pool.result(cyc, sids[j], k), alone[j][done[j]][k])]
                assert not bad, f"session {j} tick {done[j]}: {bad[:3]}"
                done[j] += 1
            times.append(pool.phase_times(cyc))
    for cyc, _ in pending:
        pool.done(cyc, wait=True)
    nph = len(stream_phases(catalog.module_ir("serve_f32")))
    assert all(len(t) == nph + 2 for t in times)
    assert all((t >= 0).all() and t[-1] > 0 for t in times)
    with pytest.raises(RuntimeError, match="not a readback buffer"):
        pool.result(cyc, sids[0], mod.getInputNames()[0])
    with pytest.raises(RuntimeError, match="replaced"):
        pool.result(1, sids[0], outs[0])
    pool.set_readback([])
    c = pool.submit([(sids[0], feeds[0][0])])
    pool.done(c, wait=True)
    assert pool.phase_times(c)[-1] < 0.05  # no readback: nothing between the last two marks


def test_tick_server_with_cycle_cap(executor):
    """The request loop ticks each session in arrival order, cycles holding
    at most half the sessions (two in flight): every tick's results equal a
    stream context's, whatever cycle it ran in."""
    from kunquant_amd.serving import TickServer
    _, mod = catalog.load("stream_f32")
    outs = mod.getOutputNames()
    sizes, T = (72, 8, 136, 64, 40), 12
    feeds = [_ticks(mod, S, T, seed=60 + j) for j, S in enumerate(sizes)]
    alone = [_alone(executor, mod, S, f, outs) for S, f in zip(sizes, feeds)]
    pool = kr.StreamPool(executor, mod, capacity_stocks=1024, max_sessions=8, max_inflight=2)
    got = {j: [] for j in range(len(sizes))}
    server = TickServer(pool, max_inflight=2,
                        on_done=lambda key, sid, t0: got[key].append({k: pool.read(sid, k) for k in outs}))
    for j, S in enumerate(sizes):
        assert server.open(j, S)
    rng = np.random.default_rng(61)
    order = [j for j in range(len(sizes)) for _ in range(T)]
    rng.shuffle(order)
    nxt = [0] * len(sizes)
    sizes_in_cycles = []
    for j in order:
        server.push(j, feeds[j][nxt[j]])
        nxt[j] += 1
        before = server.ticks_in_cycles
        server.poll()
        server.dispatch()
        if server.ticks_in_cycles > before:
            sizes_in_cycles.append(server.ticks_in_cycles - before)
    server.drain()
    assert max(sizes_in_cycles) <= 3  # ceil(5 sessions / 2 in flight)
    for j in range(len(sizes)):
        assert len(got[j]) == T
        for t in range(T):
            bad = [k for k in outs if not _same(got[j][t][k], alone[j][t][k])]
            assert not bad, f"session {j} tick {t}: {bad[:3]}"
