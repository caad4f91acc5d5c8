"""Rank of a rank on the device at the bench's universe: TsArgMax over
TsRank over TsRank over a rolling sum of the rank of volume, 8192 stocks x
1024 days, against the reference's own outputs for three stocks
(tests/golden/ref_tsrank_chain_f32.npz). The second TsRank continues the
first one's counts, as the reference's copy-constructed rank state does.

Bar: bit-exact (NaN positions equal, NaN payloads ignored)."""
import os

import numpy as np
import pytest

from kunquant_amd.runner import KunRunner as kr
from kunquant_amd.workloads import catalog, data

pytestmark = pytest.mark.gpu

GOLD = os.path.join(os.path.dirname(__file__), "golden")


def test_rank_of_rank_matches_reference():
    z = np.load(os.path.join(GOLD, "ref_tsrank_chain_f32.npz"))
    S, T, seed = (int(x) for x in z["meta"])
    st = list(z["stocks"])
    vol = data.ohlcv(S, T, seed=seed, dtype=np.float32, nan_frac=0.1)["volume"]
    _, mod = catalog.load("tsrank_chain_f32")
    got = kr.runGraph(kr.createSingleThreadExecutor(), mod, {"volume": np.ascontiguousarray(vol)}, 0, T)
    for k in ("r2", "wsum", "t60", "t30", "f022"):
        assert np.array_equal(got[k][:, st], z[f"out_{k}"], equal_nan=True), k
