"""The cross-lane wait planner of batch runs (kunquant_amd/csrc/lane_sync.hpp,
host code) on random launch DAGs: tests/native/lane_sync_check.cpp, built
with g++, follows stream order and the planned waits only and checks that
every dependency is complete before its dependent starts."""
import os
import shutil
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(os.path.dirname(HERE), "kunquant_amd", "csrc")


@pytest.mark.skipif(shutil.which("g++") is None, reason="needs g++")
def test_lane_sync_plan_covers_every_dependency(tmp_path):
    exe = str(tmp_path / "lane_sync_check")
    subprocess.run(["g++", "-O2", "-std=c++17", "-I", CSRC, "-o", exe,
                    os.path.join(HERE, "native", "lane_sync_check.cpp")], check=True)
    r = subprocess.run([exe], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stdout
    planned, naive = [int(x) for x in r.stdout.split() if x.isdigit()][:2]
    assert planned < naive  # waits are pruned
