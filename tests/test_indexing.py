"""CPU checks of the tracer's index arithmetic, compiled from the kernels' own header: the lane
groups that serve fronts of up to a quarter workgroup (tests/indexing/group_check.cpp)."""
import os
import shutil
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))


@pytest.mark.skipif(shutil.which("hipcc") is None, reason="needs hipcc (host compile only)")
def test_lane_groups_cover_every_entry_and_neighbour_once(tmp_path):
    exe = tmp_path / "group_check"
    src = os.path.join(HERE, "indexing", "group_check.cpp")
    subprocess.run(["hipcc", "-x", "hip", "--offload-arch=gfx950", "-O1", "-std=c++17", src, "-o", str(exe)],
                   check=True, capture_output=True)
    r = subprocess.run([str(exe)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0 and r.stdout.strip() == "ok", r.stdout + r.stderr
