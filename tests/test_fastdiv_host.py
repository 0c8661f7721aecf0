"""Host check of the multiply-high division the kernels use for wave-uniform block indices
(rmagine_amd/csrc/rmh_fastdiv.hpp): tests/cpp/fastdiv_check.cpp, compiled with the host
compiler, compares it with the built-in unsigned division on 3.3e7 dividend/divisor pairs
(every divisor up to 65536 around each of its multiples, powers of two and their
neighbours, the largest divisors, random pairs, and every dividend below 2^20 for typical
tile counts)."""
import shutil
import subprocess
from pathlib import Path

import pytest

SRC = Path(__file__).resolve().parent / "cpp" / "fastdiv_check.cpp"


def test_fast_div_equals_integer_division(tmp_path):
    cxx = shutil.which("g++") or shutil.which("c++")
    if cxx is None:
        pytest.skip("no host C++ compiler")
    exe = tmp_path / "fastdiv_check"
    subprocess.run([cxx, "-O2", "-std=c++17", "-Wall", "-o", str(exe), str(SRC)], check=True)
    res = subprocess.run([str(exe)], capture_output=True, text=True, timeout=120)
    assert res.returncode == 0, res.stdout + res.stderr
    assert int(res.stdout.split()[-1]) > 30_000_000
