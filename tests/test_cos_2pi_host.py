"""Host check of the Box-Muller cosine of the range-noise kernel (rmagine_amd/csrc/
rmh_noise.hip: cos_2pi): tests/cpp/cos_2pi_check.c evaluates the same operations for every
uniform the kernel can draw (2^24 values) and requires them within 1e-15 of libm's cos of the
FP64 product 2 pi u, the oracle's formula, and of the long-double cosine.  With two
quadrants exchanged the same program must fail: the check can fail."""
import shutil
import subprocess

import pytest
from pathlib import Path

SRC = Path(__file__).resolve().parent / "cpp" / "cos_2pi_check.c"


def _run(tmp_path, defines):
    cc = shutil.which("gcc") or shutil.which("cc")
    if cc is None:
        pytest.skip("no host C compiler")
    exe = tmp_path / ("cos_2pi_check" + "".join(d.replace("-D", "_") for d in defines))
    subprocess.run([cc, "-O2", "-std=c11", "-ffp-contract=off", "-Wall", *defines, "-o", str(exe), str(SRC), "-lm"],
                   check=True)
    res = subprocess.run([str(exe)], capture_output=True, text=True, timeout=120)
    n, e_libm, e_exact = res.stdout.split()
    return res.returncode, int(n), float(e_libm), float(e_exact)


def test_cos_2pi_within_1e15_of_libm_for_every_draw(tmp_path):
    rc, n, e_libm, e_exact = _run(tmp_path, [])
    assert rc == 0 and n == 1 << 24, (e_libm, e_exact)
    assert e_libm < 1e-15 and e_exact < 1e-15


def test_cos_2pi_check_fails_with_swapped_quadrants(tmp_path):
    rc, n, e_libm, e_exact = _run(tmp_path, ["-DCOS_2PI_SWAP_QUADRANTS"])
    assert rc != 0 and e_libm > 0.5
