"""Shared helpers of the test-suite."""
import numpy as np

from kunquant_amd.Driver import compileit
from kunquant_amd.codegen.program import program_text
from oracle.kun_oracle import OracleProgram, OracleStream


def oracle_module(f, name="t", **cfg):
    cfg.setdefault("input_layout", "TS")
    cfg.setdefault("output_layout", "TS")
    m = compileit(f, name, **cfg)
    return m, OracleProgram(program_text(m))


def run_oracle(f, inputs, cur_time=0, length=None, **cfg):
    m, p = oracle_module(f, **cfg)
    first = next(v for k, v in inputs.items() if not k.startswith("__init"))
    T, S = first.shape
    if length is None:
        length = T - cur_time
    return p.run(inputs, S, T, cur_time, length, nthreads=4)


def bits(a: np.ndarray) -> np.ndarray:
    return a.view(np.uint32 if a.dtype == np.float32 else np.uint64)


def assert_bit_exact(a, b, msg=""):
    a, b = np.asarray(a), np.asarray(b)
    assert a.shape == b.shape, msg
    if np.array_equal(bits(a), bits(b)):
        return
    nan_a, nan_b = np.isnan(a), np.isnan(b)
    assert np.array_equal(nan_a, nan_b), f"{msg}: NaN masks differ"
    fin = ~nan_a
    bad = np.nonzero(a[fin] != b[fin])[0]
    assert bad.size == 0, f"{msg}: {bad.size} values differ, e.g. {a[fin][bad[:3]]} vs {b[fin][bad[:3]]}"
