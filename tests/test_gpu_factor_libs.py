"""Factor libraries end to end on the device against the reference's own
libraries: this package's Alpha158 (kunquant_amd/predefined/Alpha158.py),
compiled as the reference's test_alpha158.py compiles its own (fp64,
blocking 4, partition factor 4, opt_reduce, fast log), equals the outputs
the reference's Alpha158 module computed on the same inputs
(tests/golden/ref_alpha158_f64.npz, 16 stocks x 200 days, 2% NaN cells).

Bar: bit-exact for all 158 factors (NaN positions equal, payloads ignored);
the reference's own test holds it to rtol 1e-4 / atol 1e-5 against qlib."""
import os

import numpy as np
import pytest

from kunquant_amd.runner import KunRunner as kr
from kunquant_amd.workloads import catalog

pytestmark = pytest.mark.gpu

GOLD = os.path.join(os.path.dirname(__file__), "golden")


def test_alpha158_matches_reference():
    z = np.load(os.path.join(GOLD, "ref_alpha158_f64.npz"))
    ins = {k[3:]: z[k] for k in z.files if k.startswith("in_")}
    outs = {k[4:]: z[k] for k in z.files if k.startswith("out_")}
    S, T, start, L = (int(x) for x in z["meta"])
    _, mod = catalog.load("alpha158_f64")
    assert sorted(mod.getOutputNames()) == sorted(outs)
    ex = kr.createSingleThreadExecutor()
    got = kr.runGraph(ex, mod, {k: ins[k] for k in mod.getInputNames()}, start, L)
    bad = [k for k in outs if not np.array_equal(got[k], outs[k], equal_nan=True)]
    assert not bad, f"{len(bad)} of {len(outs)} factors differ: {bad[:8]}"
