"""Float division by a constant as a double product (kun::div_const, which
the HIP emitter uses for DivConst ops of float modules) against IEEE float
division for every float x, on the host: tests/native/div_const_check.c
compiled with gcc, run for a few constants (the argument that makes the
product exact is in kun_device.hpp)."""
import os
import shutil
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))


@pytest.mark.skipif(shutil.which("gcc") is None, reason="needs gcc")
def test_div_const_matches_float_division_for_every_float(tmp_path):
    exe = str(tmp_path / "div_const_check")
    subprocess.run(["gcc", "-O2", "-o", exe, os.path.join(HERE, "native", "div_const_check.c"), "-lpthread", "-lm"],
                   check=True)
    r = subprocess.run([exe, "30", "0.1", "16777215"], capture_output=True, text=True, timeout=600)
    assert r.returncode == 0, r.stdout
    assert r.stdout.count("mismatches=0") == 3, r.stdout
