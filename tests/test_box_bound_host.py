"""Host check of the FP32 box bound of the closest-point walk (rmagine_amd/csrc/
rmh_closest.hip: box_d2): tests/cpp/box_bound_check.c computes it as the kernel does for
4e6 random queries and boxes (queries up to 1e19, gaps down to one float) and requires
that it never exceeds the squared distance to the box.  Built without the 2^-20 shrink the
same program must find violations: the check can fail."""
import shutil
import subprocess
from pathlib import Path

import pytest

SRC = Path(__file__).resolve().parent / "cpp" / "box_bound_check.c"


def _run(tmp_path, defines, n):
    cc = shutil.which("gcc") or shutil.which("cc")
    if cc is None:
        pytest.skip("no host C compiler")
    exe = tmp_path / ("box_bound_check" + "".join(d.replace("-D", "_") for d in defines))
    subprocess.run([cc, "-O2", "-std=c11", "-ffp-contract=off", "-Wall", *defines, "-o", str(exe), str(SRC), "-lm"],
                   check=True)
    res = subprocess.run([str(exe), str(n)], capture_output=True, text=True, timeout=120)
    checked, bad = map(int, res.stdout.split("\n")[-2].split())
    return res.returncode, checked, bad, res.stdout


def test_box_bound_never_exceeds_the_distance(tmp_path):
    rc, checked, bad, out = _run(tmp_path, [], 4_000_000)
    assert rc == 0 and bad == 0 and checked == 4_000_000, out[-2000:]


def test_box_bound_check_fails_without_the_shrink(tmp_path):
    rc, checked, bad, _ = _run(tmp_path, ["-DBOX_BOUND_WITHOUT_SHRINK"], 1_000_000)
    assert rc != 0 and bad > 0
