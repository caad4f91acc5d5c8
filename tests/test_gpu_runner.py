"""KunRunner.runGraph behaviour on the device: argument validation and
error messages of the reference binding (PyBinding.cpp:242-394), provided
output buffers, time windows, device-resident execution."""
import numpy as np
import pytest

from kunquant_amd.runner import KunRunner as kr
from kunquant_amd.workloads import catalog, data

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def executor():
    return kr.createSingleThreadExecutor()


def test_aligned_shape_error(executor):  # test_runtime.py:567-592
    _, mod = catalog.load("aligned_f32")
    a = np.random.default_rng(41).random((20, 24)).astype(np.float32)
    out = kr.runGraph(executor, mod, {"a": a}, 0, 20)
    import pandas as pd
    np.testing.assert_allclose(out["ou2"], pd.DataFrame(a).rank(axis=1, pct=True).to_numpy() * 2, rtol=1e-6)
    with pytest.raises(RuntimeError) as e:
        kr.runGraph(executor, mod, {"a": np.ascontiguousarray(a[:, :-3])}, 0, 20)
    assert str(e.value) == "Bad shape at a"


def test_validation_errors(executor):
    _, mod = catalog.load("simple_f32")
    d = data.ohlcv(16, 30, seed=3)
    ins = {k: d[k] for k in mod.getInputNames()}
    with pytest.raises(RuntimeError, match="Expecting float buffer"):
        kr.runGraph(executor, mod, dict(ins, close=ins["close"].astype(np.float64)), 0, 30)
    with pytest.raises(RuntimeError, match="Bad parameter: length"):
        kr.runGraph(executor, mod, ins, 0, 31)
    with pytest.raises(RuntimeError, match="Bad shape at"):
        kr.runGraph(executor, mod, dict(ins, close=np.ascontiguousarray(ins["close"][:, :8])), 0, 30)
    missing = dict(ins)
    missing.pop("close")
    with pytest.raises(RuntimeError, match="Buffer name not found: close"):
        kr.runGraph(executor, mod, missing, 0, 30)


def test_outputs_and_time_window(executor):
    _, mod = catalog.load("simple_f32")
    d = data.ohlcv(32, 60, seed=4)
    ins = {k: d[k] for k in mod.getInputNames()}
    provided = {"x": np.zeros((40, 32), np.float32)}
    out = kr.runGraph(executor, mod, ins, 20, 40, provided)
    assert out["x"] is provided["x"]
    fresh = kr.runGraph(executor, mod, {k: np.ascontiguousarray(v[20:]) for k, v in ins.items()}, 0, 40)
    for k in out:
        assert np.array_equal(out[k], fresh[k], equal_nan=True), k


def test_device_resident_matches_host_path(executor):
    import torch
    _, mod = catalog.load("cfg2_f32")
    S, T = 512, 128
    d = data.ohlcv(S, T, seed=5, nan_frac=0.1)
    host = kr.runGraph(executor, mod, d, 0, T)
    dev_in = {k: torch.from_numpy(v).cuda() for k, v in d.items()}
    dev_out = {k: torch.empty((T, S), dtype=torch.float32, device="cuda") for k in mod.getOutputNames()}
    torch.cuda.synchronize()
    kr.runGraphDevice(executor, mod, {k: v.data_ptr() for k, v in list(dev_in.items()) + list(dev_out.items())},
                      S, T, 0, T)
    executor.synchronize()
    for k, v in dev_out.items():
        assert np.array_equal(v.cpu().numpy(), host[k], equal_nan=True), k


def test_module_errors():
    lib, mod = catalog.load("simple_f32")
    with pytest.raises(RuntimeError, match="Module name not found"):
        lib.getModule("no_such_module")
    with pytest.raises(RuntimeError, match="Cannot load library"):
        kr.Library.load("/nonexistent/lib.so")
    assert mod.getOutputUnreliableCount() == {"x": 0, "std20": 20, "vsum10": 10, "rank_x": 0}


def test_device_runs_on_two_streams_share_temporaries_safely(executor):
    """The same module and shape run back to back on two different torch
    streams: both runs use the executor's cached temporaries, so the second
    must wait for the first (whatever the streams) -- both results equal the
    oracle bit for bit. The first stream is held busy by a long torch kernel
    so that, without the wait, the second run would overtake it."""
    import torch
    from kunquant_amd.codegen.program import program_text
    from oracle.kun_oracle import OracleProgram
    _, mod = catalog.load("cfg2_f32")
    S, T = 640, 160
    runs = []
    for seed in (11, 12):
        d = data.ohlcv(S, T, seed=seed, nan_frac=0.1)
        dev_in = {k: torch.from_numpy(v).cuda() for k, v in d.items()}
        dev_out = {k: torch.full((T, S), 7.0, dtype=torch.float32, device="cuda") for k in mod.getOutputNames()}
        runs.append((d, dev_in, dev_out))
    torch.cuda.synchronize()
    streams = [torch.cuda.Stream(), torch.cuda.Stream()]
    big = torch.rand(4096, 4096, device="cuda")
    streams[0].wait_stream(torch.cuda.current_stream())
    with torch.cuda.stream(streams[0]):
        for _ in range(8):
            big = big @ big * 1e-3   # keep stream 0 busy before its graph run
    for (d, dev_in, dev_out), st in zip(runs, streams):
        ptrs = {k: v.data_ptr() for k, v in list(dev_in.items()) + list(dev_out.items())}
        kr.runGraphDevice(executor, mod, ptrs, S, T, 0, T, stream=st.cuda_stream)
    torch.cuda.synchronize()
    prog = OracleProgram(program_text(catalog.module_ir("cfg2_f32")))
    for i, (d, _, dev_out) in enumerate(runs):
        exp = prog.run({k: d[k] for k in mod.getInputNames()}, S, T, 0, T, nthreads=16)
        bad = [k for k in exp if not np.array_equal(dev_out[k].cpu().numpy(), exp[k], equal_nan=True)]
        assert not bad, f"run {i} on its own stream: outputs {bad[:3]} differ from the oracle"


def test_autotune_cache(executor, tmp_path):
    """A pick stored in the cache file is applied by a later tune() without
    timing anything; results do not depend on the policy."""
    import torch
    from kunquant_amd import autotune
    _, mod = catalog.load("simple_f32")
    S, T = 256, 200
    d = data.ohlcv(S, T, seed=9)
    dev = {k: torch.from_numpy(d[k]).cuda() for k in mod.getInputNames()}
    out = {k: torch.empty((T, S), dtype=torch.float32, device="cuda") for k in mod.getOutputNames()}
    ptrs = {k: v.data_ptr() for k, v in list(dev.items()) + list(out.items())}
    cache = str(tmp_path / "tune.json")
    best, times = autotune.tune(executor, mod, ptrs, S, T, reps=1, cache=cache)
    assert times and best in times
    first = {k: v.cpu().numpy() for k, v in out.items()}
    again, t2 = autotune.tune(executor, mod, ptrs, S, T, reps=1, cache=cache)
    assert again == best and t2 == {}
    kr.runGraphDevice(executor, mod, ptrs, S, T, 0, T)
    executor.synchronize()
    for k, v in out.items():
        assert np.array_equal(v.cpu().numpy(), first[k], equal_nan=True), k
