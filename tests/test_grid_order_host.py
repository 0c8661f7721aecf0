"""Host check of the grid tile order of O1Dn/OnDn sensors (rmagine_amd/csrc/rmh_grid_order.hpp,
used by rmh_sensor_set_ray_order with RMH_RAY_ORDER_TILES): tests/cpp/grid_order_check.cpp,
compiled with the host compiler, checks that the order is a permutation of the buffer
indices for every grid up to 70 x 21 and a few scanner sizes, and that each run of 64
slots is one 4 x 16 block when the grid's sides allow it."""
import shutil
import subprocess
from pathlib import Path

import pytest

SRC = Path(__file__).resolve().parent / "cpp" / "grid_order_check.cpp"


def test_grid_tile_order_is_a_tiled_permutation(tmp_path):
    cxx = shutil.which("g++") or shutil.which("c++")
    if cxx is None:
        pytest.skip("no host C++ compiler")
    exe = tmp_path / "grid_order_check"
    subprocess.run([cxx, "-O2", "-std=c++17", "-Wall", "-o", str(exe), str(SRC)], check=True)
    res = subprocess.run([str(exe)], capture_output=True, text=True, timeout=120)
    assert res.returncode == 0, res.stdout + res.stderr
    assert int(res.stdout.split()[-1]) == 70 * 21 + 6
