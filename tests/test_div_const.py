"""Float division by a constant as a double product (kun::div_const, which
the HIP emitter uses for DivConst ops of float modules) against IEEE float
division for every float x, on the host: tests/native/div_const_check.c
compiled with gcc, run for a few constants; and the bare product, which the
emitter uses where its per-divisor check (div_product_exact) finds every
exact tie between subnormal floats rounded right, against that check (the
argument that only such ties can go wrong is in kun_device.hpp)."""
import os
import shutil
import subprocess

import pytest

from kunquant_amd.codegen.hip import div_product_exact

HERE = os.path.dirname(os.path.abspath(__file__))


@pytest.mark.skipif(shutil.which("gcc") is None, reason="needs gcc")
def test_div_const_matches_float_division_for_every_float(tmp_path):
    exe = str(tmp_path / "div_const_check")
    subprocess.run(["gcc", "-O2", "-o", exe, os.path.join(HERE, "native", "div_const_check.c"), "-lpthread", "-lm"],
                   check=True)
    # kun::div_rc (subnormal results divide) for every float x
    consts = ["30", "0.1", "16777215", "16350"]
    r = subprocess.run([exe] + consts, capture_output=True, text=True, timeout=900)
    assert r.returncode == 0, r.stdout
    assert r.stdout.count("mismatches=0") == len(consts), r.stdout
    # the bare product (kun::div_const): wrong for 16350 = 2 * 8175 and
    # 32614 = 2 * 16307, where x / c is exactly a midpoint between subnormal
    # floats for some x; the emitter's per-divisor check agrees with the
    # exhaustive one, also for divisors with midpoints that round right
    bare = ["16350", "32614", "30", "10", "6"]
    r = subprocess.run([exe, "-bare"] + bare, capture_output=True, text=True, timeout=900)
    for c in bare:
        exact = f"c={float(c):.9g} mismatches=0" in r.stdout.splitlines()
        assert exact == div_product_exact(float(c)), (c, r.stdout)
    assert not div_product_exact(16350.0) and div_product_exact(30.0)
