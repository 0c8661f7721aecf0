"""Host check that the ray kernels' FP32 slab test is conservative (rmagine_amd/csrc/rmh_trace.hpp:
trace()'s SlabRay set-up and slab2): tests/cpp/slab_check.c builds 4e6 random rays and boxes
that the exact ray meets at some t* in [0, tbf] (degenerate and grazing boxes, origins up to
1e4, tiny and zero direction components, tbf = range_max or t* rounded up after a hit) and
requires that the FP32 test, computed as the kernels compute it, culls none of them and never
reports an entry distance beyond t*.  The same program built without the range term of the
per-ray margin must cull some: the check can fail."""
import shutil
import subprocess

import pytest
from pathlib import Path

SRC = Path(__file__).resolve().parent / "cpp" / "slab_check.c"


def _run(tmp_path, defines, n):
    cc = shutil.which("gcc") or shutil.which("cc")
    if cc is None:
        pytest.skip("no host C compiler")
    exe = tmp_path / ("slab_check" + "".join(d.replace("-D", "_") for d in defines))
    subprocess.run([cc, "-O2", "-std=c11", "-Wall", *defines, "-o", str(exe), str(SRC), "-lm"], check=True)
    res = subprocess.run([str(exe), str(n)], capture_output=True, text=True, timeout=120)
    checked, culled = map(int, res.stdout.split("\n")[-2].split())
    return res.returncode, checked, culled, res.stdout


def test_slab_test_never_culls_a_box_the_ray_meets(tmp_path):
    rc, checked, culled, out = _run(tmp_path, [], 4_000_000)
    assert rc == 0 and culled == 0, out[-2000:]
    assert checked == 4_000_000


def test_slab_check_fails_without_the_range_term(tmp_path):
    rc, checked, culled, _ = _run(tmp_path, ["-DSLAB_WITHOUT_RANGE_TERM"], 1_000_000)
    assert rc != 0 and culled > 0
