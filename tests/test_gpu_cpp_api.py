"""Runs the C++ drop-in test programs (tests/cpp) on the GPU; they throw on failure."""
import subprocess
from pathlib import Path

import pytest

pytestmark = pytest.mark.gpu

BUILD = Path(__file__).resolve().parent / "cpp" / "_build"


@pytest.mark.parametrize("prog", ["simulation_kat", "correction_rcc"])
def test_cpp_program(prog):
    exe = BUILD / prog
    assert exe.exists(), f"{exe} not built (run __graft_entry__.build())"
    maps = Path(__file__).resolve().parent / "golden" / "maps"
    res = subprocess.run([str(exe), str(maps)], capture_output=True, text=True, timeout=120)
    print(res.stdout, res.stderr)
    assert res.returncode == 0, res.stdout + res.stderr
