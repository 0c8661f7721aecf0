"""Host check of the BVH topology code the device builder runs (rmagine_amd/csrc/rmh_treelet.hpp):
tests/cpp/treelet_check.cpp is compiled with the host compiler and run on meshes of the
benchmark families.  After the Karras tree and after every treelet round it verifies that
the arrays form one binary tree over all triangles (every node reached once, parent links,
boxes = unions, counts = sums), that the SAH cost does not rise, and that the depth-first
leaf numbering makes every subtree one contiguous range."""
import json
import shutil
import subprocess
from pathlib import Path

import numpy as np
import pytest

from rmagine_amd import scenes

SRC = Path(__file__).resolve().parent / "cpp" / "treelet_check.cpp"


@pytest.fixture(scope="module")
def checker(tmp_path_factory):
    cxx = shutil.which("g++") or shutil.which("c++")
    if cxx is None:
        pytest.skip("no host C++ compiler")
    exe = tmp_path_factory.mktemp("treelet") / "treelet_check"
    subprocess.run([cxx, "-O2", "-std=c++17", "-Wall", "-o", str(exe), str(SRC)], check=True)
    return exe


def _run(exe, tmp_path, V, F, rounds=3):
    V = np.asarray(V, np.float32)
    F = np.asarray(F, np.uint32)
    p = tmp_path / "mesh.bin"
    with open(p, "wb") as f:
        np.array([len(V), len(F)], np.uint32).tofile(f)
        V.tofile(f)
        F.tofile(f)
    res = subprocess.run([str(exe), str(p), str(rounds)], capture_output=True, text=True, timeout=300)
    assert res.returncode == 0, res.stdout + res.stderr
    return json.loads(res.stdout)


def test_room_and_terrain(checker, tmp_path):
    V, F, _ = scenes.room_workload(n_sensors=1)
    r = _run(checker, tmp_path, V, F)
    assert r["faces"] == len(F) and r["restructured"] > 0 and r["sah_treelets"] < 0.9 * r["sah_karras"]
    V, F, _ = scenes.terrain_workload(n_grid=128, n_sensors=1)
    r = _run(checker, tmp_path, V, F)
    assert r["sah_treelets"] < r["sah_karras"]


def test_objects_and_degenerate(checker, tmp_path):
    V, F = scenes.objects_mesh(n_objects=300, n_ico=80)
    r = _run(checker, tmp_path, V, F)
    assert r["sah_treelets"] < r["sah_karras"]
    # many coincident triangles (equal Morton keys everywhere) and a two-triangle mesh
    V = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0]], np.float32)
    F = np.zeros((500, 3), np.uint32)
    F[:] = [0, 1, 2]
    _run(checker, tmp_path, V, F)
    _run(checker, tmp_path, V, F[:2])
