"""Batched uneven jobs on the device: every job's result is the same when it
runs alone (plain runGraph), in a batch of jobs enqueued together, and in
the batch captured as one HIP graph and replayed; the chunking autotuner
changes no result. Each job is also checked against the oracle.

Bar: bit-exact (NaN positions equal, NaN payload bits ignored)."""
import numpy as np
import pytest

from kunquant_amd import autotune
from kunquant_amd.codegen.program import program_text
from kunquant_amd.runner import KunRunner as kr
from kunquant_amd.workloads import catalog, data, jobs
from oracle.kun_oracle import OracleProgram

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


def _same(a, b):
    return np.array_equal(a, b, equal_nan=True)


@pytest.fixture(scope="module")
def batch():
    specs = jobs.draw_jobs(n=10, seed=7, stocks=(50, 900), days=(250, 420))
    mods = {g: catalog.load(g)[1] for g in jobs.JOB_GRAPHS}
    ex = kr.createSingleThreadExecutor()
    items = []
    for sp in specs:
        mod = mods[sp.graph]
        d = data.ohlcv(sp.stocks, sp.days, seed=sp.seed, nan_frac=0.1)
        ins = {k: d[k] for k in mod.getInputNames()}
        dev_in = {k: torch.from_numpy(v).to("cuda:0") for k, v in ins.items()}
        dev_out = {k: torch.empty((sp.days, sp.stocks), dtype=torch.float32, device="cuda:0")
                   for k in mod.getOutputNames()}
        items.append((sp, mod, ins, dev_in, dev_out))
    alone = [kr.runGraph(ex, mod, ins, 0, sp.days) for sp, mod, ins, _, _ in items]
    return ex, items, alone


def _job_tuples(items):
    return [(mod, {k: v.data_ptr() for k, v in list(di.items()) + list(do.items())}, sp.stocks, sp.days, 0,
             sp.days) for sp, mod, _, di, do in items]


def _fetch(items):
    torch.cuda.synchronize()
    return [{k: v.cpu().numpy() for k, v in do.items()} for _, _, _, _, do in items]


def _poison(items):
    """Overwrite every output (torch's stream), then wait for it: the
    executor enqueues on a stream of its own."""
    for *_, do in items:
        for v in do.values():
            v.fill_(12345.0)
    torch.cuda.synchronize()


def _check(got, alone, what):
    for j, (g, a) in enumerate(zip(got, alone)):
        bad = [k for k in a if not _same(g[k], a[k])]
        assert not bad, f"{what}: job {j} outputs {bad[:3]} differ from the job run alone"


def test_jobs_alone_match_oracle(batch):
    _, items, alone = batch
    progs = {}
    for (sp, mod, ins, _, _), got in zip(items, alone):
        if sp.graph not in progs:
            progs[sp.graph] = OracleProgram(program_text(catalog.module_ir(sp.graph)))
        exp = progs[sp.graph].run(ins, sp.stocks, sp.days, 0, sp.days, nthreads=16)
        bad = [k for k in exp if not _same(got[k], exp[k])]
        assert not bad, f"{sp}: outputs {bad[:3]} differ from the oracle"


def test_batched_jobs_match_alone(batch):
    ex, items, alone = batch
    _poison(items)
    kr.runGraphBatchDevice(ex, _job_tuples(items))
    ex.synchronize()
    _check(_fetch(items), alone, "batched")
    # again, with the executor's cached temporaries reused
    _poison(items)
    kr.runGraphBatchDevice(ex, _job_tuples(items))
    ex.synchronize()
    _check(_fetch(items), alone, "batched (second call)")


def test_batch_graph_replays_match_alone(batch):
    """The batch as one HIP graph (built node by node from the launch DAGs),
    replayed three times, with eager device work (a plain device run of one
    job into other buffers, torch fills of every output) between the
    replays; every replay is checked."""
    ex, items, alone = batch
    g = kr.BatchGraph(ex, _job_tuples(items))
    sp0, mod0, _, di0, do0 = items[0]
    side = {k: torch.empty_like(v) for k, v in do0.items()}
    side_ptrs = {k: v.data_ptr() for k, v in list(di0.items()) + list(side.items())}
    for rep in range(3):
        _poison(items)
        g.launch()
        ex.synchronize()
        _check(_fetch(items), alone, f"graph replay {rep}")
        kr.runGraphDevice(ex, mod0, side_ptrs, sp0.stocks, sp0.days, 0, sp0.days)
        ex.synchronize()
        got = {k: v.cpu().numpy() for k, v in side.items()}
        assert all(_same(got[k], alone[0][k]) for k in got), f"eager run between replays {rep}"
    del g


def test_autotune_keeps_results(batch):
    ex, items, alone = batch
    j = max(range(len(items)), key=lambda i: items[i][0].stocks * items[i][0].days)
    sp, mod, _, di, do = items[j]
    ptrs = {k: v.data_ptr() for k, v in list(di.items()) + list(do.items())}
    best, times = autotune.tune(ex, mod, ptrs, sp.stocks, sp.days, reps=1,
                                candidates=((4096, 3, 64), (65536, 2, 16), (32768, 3, 64)))
    assert best in times and len(times) == 3
    for v in do.values():
        v.fill_(-1.0)
    torch.cuda.synchronize()
    kr.runGraphDevice(ex, mod, ptrs, sp.stocks, sp.days, 0, sp.days)
    ex.synchronize()
    got = {k: v.cpu().numpy() for k, v in do.items()}
    bad = [k for k in got if not _same(got[k], alone[j][k])]
    assert not bad, f"tuned chunking {best} changed outputs {bad[:3]}"

