"""Generated stage kernels run on the CPU (g++ build of the same HIP source,
lane by lane) against the oracle: every time-series stage's outputs, from
the oracle's values of its inputs, equal the oracle's bit for bit (NaN
positions equal), for one chunk, one-row chunks and ragged chunks."""
import numpy as np
import pytest

from kunquant_amd.codegen.program import program_text
from kunquant_amd.workloads import catalog, data
from oracle.kun_oracle import OracleProgram
from emu.emu import EmuModule

MODULES = ["simple_f32", "ops_f32", "ops_f64", "ops_f32_sts", "aligned_f32", "rand40_f64",
           "streamgraph_batch_f64", "alpha158_f64"]


def _check_module(name, S, T, start, L, chunks, seed=3, nan_frac=0.05, lds_budget=None):
    ir = catalog.module_ir(name)
    emu = EmuModule(ir, lds_budget)
    dt = np.float64 if ir.dtype == "double" else np.float32
    d = data.ohlcv(S, T, seed=seed, dtype=dt, nan_frac=nan_frac)
    prog = OracleProgram(program_text(ir))
    in_names = [b.name for b in ir.buffers if b.kind == "INPUT"]
    ins = {}
    for n in in_names:
        x = d[n] if n in d else d["close"]
        if ir.input_layout == "STs":
            b = ir.blocking_len
            Sp = (S + b - 1) // b * b
            xp = np.zeros((T, Sp), dtype=dt)
            xp[:, :S] = x
            x = np.ascontiguousarray(xp.reshape(T, Sp // b, b).transpose(1, 0, 2))
        ins[n] = np.ascontiguousarray(x)
    exp = prog.run(ins, S, T, start, L, nthreads=4, temps=True)
    bufs = dict(ins)
    bufs.update(exp)
    failures = []
    for st in ir.stages:
        if st.kind != "ts":
            continue
        for chunk in chunks:
            got = dict(bufs)
            for b in st.out_bufs:
                got[b.name] = np.full_like(exp[b.name], 12345.0)
            emu.run_stage(st.idx, got, S, T, start, L, chunk)
            for b in st.out_bufs:
                g, e = got[b.name], exp[b.name]
                if g.ndim == 3:  # STs output: padding lanes past S are not computed
                    g = g.transpose(1, 0, 2).reshape(L, -1)[:, :S]
                    e = e.transpose(1, 0, 2).reshape(L, -1)[:, :S]
                if not np.array_equal(g, e, equal_nan=True):
                    failures.append((st.idx, chunk, b.name))
    assert not failures, f"{name}: stages/chunks/buffers differing: {failures[:10]}"


@pytest.mark.parametrize("name", MODULES)
def test_emulated_stages_match_oracle(name):
    _check_module(name, S=70, T=140, start=6, L=130, chunks=(0, 1, 17))


@pytest.mark.parametrize("name", ["ops_f32", "ops_f64", "ops_f32_sts", "rand40_f64"])
def test_emulated_stages_without_lds(name):
    """A 1 KiB LDS budget moves history rings of computed values to HBM
    scratch and makes window loops re-read HBM values directly."""
    _check_module(name, S=70, T=140, start=6, L=130, chunks=(0, 11), lds_budget=1024)


def test_emulated_config2_stages_match_oracle():
    _check_module("cfg2_f32", S=20, T=200, start=0, L=200, chunks=(0, 9), seed=4)
