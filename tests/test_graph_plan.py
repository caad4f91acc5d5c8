"""Host-side checks (no GPU) of the HIP graph a batch of jobs becomes, and
of the module-cache keys of the runtime.

The graph is built node by node from the jobs' launch DAGs by the same code
that builds it on the device, here with a recording builder in place of
hipGraphAddKernelNode (kunDescribeGraphBatch). Checked: every node depends
only on nodes added before it (so the graph is acyclic and every handle
exists when used); the nodes of one job depend only on that job's nodes;
each launch-DAG node maps to a chain of kernels (two for a Scale stage: row
sums, then the division); the first kernel of a DAG node depends on exactly
the last kernel of each of its DAG dependencies and of the DAG node before it
on the same lane (the arena's memory plan reuses slots along lane order, as
the executor's streams serialise it), later kernels on the one before them;
every DAG node is present once."""
import ctypes
from collections import defaultdict

import pytest

from kunquant_amd.runner import KunRunner as kr
from kunquant_amd.workloads import catalog, jobs as wjobs


def parse(text):
    nodes, dag = {}, defaultdict(dict)
    for line in text.strip().splitlines():
        head, deps = line.split(":")
        deps = [int(x) for x in deps.split()]
        f = head.split()
        if f[0] == "dag":
            job, n, stage, kern, lane = map(int, f[1:6])
            dag[job][n] = (stage, kern, lane, deps)
        else:
            nid, kind, job, stage, kern, grid, block = int(f[0]), f[1], *map(int, f[2:7])
            nodes[nid] = dict(kind=kind, job=job, stage=stage, kern=kern, grid=grid, block=block, deps=deps)
    return nodes, dag


def check_graph(text, mods):
    nodes, dag = parse(text)
    ids = sorted(nodes)
    assert ids == list(range(len(ids)))
    for i in ids:
        for d in nodes[i]["deps"]:
            assert d < i, f"node {i} depends on {d}, added later"
            assert nodes[d]["job"] == nodes[i]["job"], f"node {i} depends on another job's node {d}"
    # nodes are added DAG node by DAG node: split each job's nodes into the
    # chains of consecutive nodes that carry the same (stage, kernel)
    chains = defaultdict(list)
    cur = None
    for i in ids:
        nd = nodes[i]
        key = (nd["job"], nd["stage"], nd["kern"])
        if cur is None or cur[0] != key or (nd["deps"] != [cur[1][-1]]):
            cur = (key, [i])
            chains[nd["job"]].append(cur)
        else:
            cur[1].append(i)
    for job, dnodes in dag.items():
        mod = mods[job]
        kinds = {i: st[1] for i, st in enumerate(mod.stages())}
        got = chains[job]
        assert len(got) == len(dnodes), f"job {job}: {len(got)} node chains for {len(dnodes)} DAG nodes"
        last_of, prev_on_lane = {}, {}
        for n, (key, chain) in enumerate(got):
            stage, kern, lane, deps = dnodes[n]
            assert key == (job, stage, kern), f"job {job} DAG node {n}: graph chain {key}"
            preds = set(deps)
            if lane in prev_on_lane:
                preds.add(prev_on_lane[lane])
            prev_on_lane[lane] = n
            want = sorted(last_of[d] for d in preds)
            assert sorted(nodes[chain[0]]["deps"]) == want, f"job {job} DAG node {n}: deps"
            for a, b in zip(chain, chain[1:]):
                assert nodes[b]["deps"] == [a]
            # kernels per DAG node: stage kernel / rank 1, scale 2 (sums, division)
            expect = 2 if kinds[stage] == 2 else 1
            assert len(chain) == expect, f"job {job} stage {stage}: {len(chain)} kernels"
            assert all(nodes[i]["kind"] == "kernel" for i in chain)
            last_of[n] = chain[-1]
    return nodes, dag


@pytest.mark.parametrize("lanes", [1, 6])
def test_batch_graph_follows_launch_dags(lanes):
    specs = wjobs.draw_jobs(n=10, seed=7, stocks=(50, 900), days=(250, 420))
    mods = [catalog.load(sp.graph)[1] for sp in specs]
    text = kr.describeGraphBatch([(m, sp.stocks, sp.days, 0, sp.days) for m, sp in zip(mods, specs)], lanes)
    nodes, dag = check_graph(text, mods)
    assert len(dag) == len(specs)
    assert sum(len(d) for d in dag.values()) <= len(nodes)


def test_batch_graph_of_the_100_factor_module():
    _, mod = catalog.load("cfg2_f32")
    text = kr.describeGraphBatch([(mod, 8192, 1024, 0, 1024), (mod, 700, 300, 20, 280)])
    nodes, dag = check_graph(text, [mod, mod])
    ranks = sum(1 for st in mod.stages() if st[1] == 1)
    assert sum(1 for n in nodes.values() if n["job"] == 0 and n["grid"] == 1024 and n["block"] == 1024) >= ranks


def test_describe_errors():
    _, mod = catalog.load("simple_f32")
    with pytest.raises(RuntimeError, match="cur_time"):
        kr.describeGraphBatch([(mod, 64, 100, 90, 20)])


def test_module_cache_key_changes_with_each_load():
    """Caches of the runtime key by (descriptor, load serial): a library
    unloaded and loaded again (its descriptors possibly at the same
    addresses) gets a new serial, so stale cache entries never match."""
    L = kr._load()
    path = catalog.library_path("simple_f32").encode()
    serials, addrs = [], []
    for _ in range(3):
        lib = L.kunLoadLibrary(path)
        m = L.kunGetModuleFromLibrary(lib, b"simple_f32")
        assert m
        serials.append(L.kunModuleLoadSerial(m))
        addrs.append(m)
        L.kunUnloadLibrary(lib)
    assert len(set(serials)) == 3 and min(serials) > 0
    lib1 = L.kunLoadLibrary(path)
    lib2 = L.kunLoadLibrary(catalog.library_path("cs_f32").encode())
    a = L.kunGetModuleFromLibrary(lib1, b"simple_f32")
    b = L.kunGetModuleFromLibrary(lib2, b"cs_f32")
    assert L.kunModuleLoadSerial(a) != L.kunModuleLoadSerial(b)
    L.kunUnloadLibrary(lib2)
    L.kunUnloadLibrary(lib1)
