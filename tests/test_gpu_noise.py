"""rmh_noise_apply on the GPU against the noise oracle (same Philox words; the FP64
transcendentals may differ in the last bit between device and host libraries, so ranges are
compared to 1 float ulp and dust decisions exactly up to a handful of boundary cases)."""
import numpy as np
import pytest
import torch

from oracle import noise as N
from rmagine_amd import noise

pytestmark = pytest.mark.gpu


def _ranges(n, seed=0):
    rng = np.random.default_rng(seed)
    r = rng.uniform(0.1, 60.0, n).astype(np.float32)
    r[::97] = 101.0   # misses (range_max + 1)
    return r


@pytest.mark.parametrize("cls,args,kind", [
    (noise.GaussianNoiseHip, (0.01, 0.05), 0),
    (noise.RelGaussianNoiseHip, (0.0, 0.02, 0.7), 1),
    (noise.RelGaussianNoiseHip, (0.005, 0.01, 1.0), 1),
    (noise.UniformDustNoiseHip, (0.01, 0.95), 2),
])
def test_matches_oracle_over_calls(cls, args, kind):
    n = 1_000_003
    r = _ranges(n)
    opts = noise.Options(seed=1234, max_range=80.0)
    nz = cls(*args, options=opts)
    t = torch.as_tensor(r, device="cuda")
    want = r.copy()
    for epoch in range(3):  # successive calls draw new numbers
        nz.apply(t)
        p = list(args) + [0.0] * (3 - len(args))
        want = N.apply(want, kind, p[0], p[1], p[2], max_range=80.0, seed=1234, epoch=epoch)
        got = t.cpu().numpy()
        if kind == 2:
            assert (got != want).sum() <= 2, (got != want).sum()
        else:
            ulp = np.spacing(np.abs(want))
            assert (np.abs(got.astype(np.float64) - want) <= ulp).all()
    if kind != 2:
        np.testing.assert_array_equal(got[r > 80.0], r[r > 80.0])  # above max_range untouched


def test_empty_and_errors():
    nz = noise.GaussianNoiseHip(0.0, 1.0)
    nz.apply(torch.empty(0, device="cuda"))
    with pytest.raises(ValueError):
        nz.apply(torch.zeros(4, dtype=torch.float64, device="cuda"))
