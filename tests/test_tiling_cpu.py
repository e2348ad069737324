"""The column tiling of the z-marching kernels (csrc/dmx_tiling.h),
compiled for the host with g++ by tests/tiling_check.cpp: every column and
plane of many domain sizes lies in exactly one tile within the tile
limits, and a 256^3 domain's 257 faces a row take seven tiles of 32 and one
of 33."""
import os
import shutil
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "..", "dumses-hybrid_amd", "csrc")


@pytest.mark.skipif(shutil.which("g++") is None, reason="no host C++ compiler")
def test_tiling_covers_every_column_once(tmp_path):
    exe = str(tmp_path / "tiling_check")
    subprocess.run(["g++", "-O2", "-std=c++17", "-I", CSRC, os.path.join(HERE, "tiling_check.cpp"), "-o", exe],
                   check=True, capture_output=True, text=True)
    r = subprocess.run([exe], capture_output=True, text=True, timeout=120)
    print(r.stdout)
    assert r.returncode == 0, r.stdout
    assert "bad 0" in r.stdout
    assert "8 x strips (7 of 32, 1 of 33), 268 tiles per block of planes, 16 blocks of planes" in r.stdout
