"""Noise oracle: Philox4x32-10 pinned by the Random123 known-answer vectors, the noise
models by their closed-form statistics (reference semantics: src/rmagine_core/src/noise/*.cpp)."""
import numpy as np
import pytest

from oracle import noise as N

# Random123 kat_vectors, philox4x32 with 10 rounds: counter, key, expected output
KAT = [((0, 0, 0, 0), (0, 0), (0x6627E8D5, 0xE169C58D, 0xBC57AC4C, 0x9B00DBD8)),
       ((0xFFFFFFFF,) * 4, (0xFFFFFFFF, 0xFFFFFFFF), (0x408F276D, 0x41C83B0E, 0xA20BC7C6, 0x6D5451FD)),
       ((0x243F6A88, 0x85A308D3, 0x13198A2E, 0x03707344), (0xA4093822, 0x299F31D0),
        (0xD16CFE09, 0x94FDCCEB, 0x5001E420, 0x24126EA1))]


@pytest.mark.parametrize("ctr,key,want", KAT)
def test_philox_known_answers(ctr, key, want):
    got = N.philox4x32_10(np.array(ctr, np.uint64), key[0], key[1])
    assert [int(x) for x in got] == list(want)


def test_gaussian_moments_and_max_range():
    n = 400_000
    r = np.full(n, 5.0, np.float32)
    r[:1000] = 500.0  # above max_range: untouched
    out = N.apply(r, 0, 0.02, 0.1, max_range=100.0, seed=7, epoch=3)
    np.testing.assert_array_equal(out[:1000], 500.0)
    d = out[1000:].astype(np.float64) - 5.0
    assert abs(d.mean() - 0.02) < 4 * 0.1 / np.sqrt(n)
    assert abs(d.std() - 0.1) < 0.002
    # a normal: ~68.3 % within one sigma, ~95.4 % within two
    z = (d - 0.02) / 0.1
    assert abs((np.abs(z) < 1).mean() - 0.6827) < 0.004 and abs((np.abs(z) < 2).mean() - 0.9545) < 0.003


def test_rel_gaussian_scales_with_range():
    n = 200_000
    for rng_val in (1.0, 4.0, 16.0):
        d = N.apply(np.full(n, rng_val, np.float32), 1, 0.0, 0.01, 0.5, seed=1).astype(np.float64) - rng_val
        assert abs(d.std() - 0.01 * rng_val ** 0.5) < 0.01 * rng_val ** 0.5 * 0.01


def test_uniform_dust_probabilities():
    n = 400_000
    for r0, hit, ret in ((10.0, 0.02, 0.9), (50.0, 0.01, 0.99), (5.0, 0.0, 1.0)):
        out = N.apply(np.full(n, r0, np.float32), 2, hit, ret, seed=3)
        changed = out != np.float32(r0)
        # P(change) = (1 - (1-h)^d) * E_u[ret^(u d)] = (1 - (1-h)^d) * (ret^d - 1) / (d ln ret)
        p_hit = 1 - (1 - hit) ** r0
        e_ret = 1.0 if ret == 1.0 else (ret ** r0 - 1) / (r0 * np.log(ret))
        want = p_hit * e_ret
        assert abs(changed.mean() - want) < 5 * np.sqrt(want * (1 - want) / n) + 1e-9
        assert (out[changed] < r0).all() and (out[changed] >= 0).all()


def test_seed_and_epoch_give_new_numbers():
    r = np.full(1000, 3.0, np.float32)
    a = N.apply(r, 0, 0.0, 1.0, seed=1, epoch=0)
    b = N.apply(r, 0, 0.0, 1.0, seed=1, epoch=1)
    c = N.apply(r, 0, 0.0, 1.0, seed=2, epoch=0)
    np.testing.assert_array_equal(a, N.apply(r, 0, 0.0, 1.0, seed=1, epoch=0))
    assert (a != b).mean() > 0.99 and (a != c).mean() > 0.99
    assert abs(np.corrcoef(a, b)[0, 1]) < 0.1
