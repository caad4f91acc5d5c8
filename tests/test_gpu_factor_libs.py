"""Factor libraries end to end on the device against the reference's own
libraries:

* this package's Alpha158 (kunquant_amd/predefined/Alpha158.py), compiled
  as the reference's test_alpha158.py compiles its own (fp64, blocking 4,
  partition factor 4, opt_reduce, fast log), equals the outputs the
  reference's Alpha158 module computed on the same inputs
  (tests/golden/ref_alpha158_f64.npz, 16 stocks x 200 days, 2% NaN cells);
* this package's Alpha101 (kunquant_amd/predefined/Alpha101.py, 82
  factors), compiled as the reference's test_alpha101.py:27-47 compiles its
  own (opt_reduce, fast log; fp32, and fp64 with fast statistics), equals
  the reference's Alpha101 module on the same inputs
  (tests/golden/ref_alpha101_f*.npz: 20 stocks x 320 days with flat days,
  the last 64 rows stored).

Bar: bit-exact for every factor (NaN positions equal, payloads ignored);
the reference's own tests hold these libraries to rtol 1e-4 / atol 1e-5
(Alpha158 against qlib) and to test_alpha101.py:155-254's per-alpha
tolerances (Alpha101 against pandas)."""
import os

import numpy as np
import pytest

from kunquant_amd.runner import KunRunner as kr
from kunquant_amd.workloads import catalog

pytestmark = pytest.mark.gpu

GOLD = os.path.join(os.path.dirname(__file__), "golden")


@pytest.mark.parametrize("golden,module", [("alpha158_f64", "alpha158_f64"), ("alpha101_f32", "alpha101_f32"),
                                           ("alpha101_f64", "alpha101_f64")])
def test_factor_library_matches_reference(golden, module):
    z = np.load(os.path.join(GOLD, f"ref_{golden}.npz"))
    ins = {k[3:]: z[k] for k in z.files if k.startswith("in_")}
    outs = {k[4:]: z[k] for k in z.files if k.startswith("out_")}
    meta = [int(x) for x in z["meta"]]
    S, T, start, L = meta[:4]
    k0 = meta[4] if len(meta) > 4 else 0   # first stored output row
    _, mod = catalog.load(module)
    assert sorted(mod.getOutputNames()) == sorted(outs)
    ex = kr.createSingleThreadExecutor()
    got = kr.runGraph(ex, mod, {k: ins[k] for k in mod.getInputNames()}, start, L)
    bad = [k for k in outs if not np.array_equal(got[k][k0:], outs[k], equal_nan=True)]
    assert not bad, f"{len(bad)} of {len(outs)} factors differ: {bad[:8]}"
    # the stored rows hold values, not only warm-up NaNs
    assert sum(np.isfinite(v).any() for v in outs.values()) >= 0.9 * len(outs)


def test_alpha101_through_the_kunquant_alias():
    """Code written against the reference's package names (the KunQuant
    alias of kunquant_amd/compat): Library.load / getModule / runGraph of
    the Alpha101 module give the reference's outputs."""
    import importlib
    import sys
    compat = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "kunquant_amd", "compat")
    sys.path.insert(0, compat)
    try:
        kr = importlib.import_module("KunQuant.runner.KunRunner")
        z = np.load(os.path.join(GOLD, "ref_alpha101_f32.npz"))
        ins = {k[3:]: z[k] for k in z.files if k.startswith("in_")}
        meta = [int(x) for x in z["meta"]]
        lib = kr.Library.load(catalog.library_path("alpha101_f32"))
        mod = lib.getModule("alpha101_f32")
        executor = kr.createMultiThreadExecutor(8)
        out = kr.runGraph(executor, mod, ins, meta[2], meta[3])
        bad = [k for k in z.files if k.startswith("out_")
               and not np.array_equal(out[k[4:]][meta[4]:], z[k], equal_nan=True)]
        assert not bad, bad[:5]
    finally:
        sys.path.remove(compat)
