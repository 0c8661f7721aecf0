"""Pins the FP64 statistics / Umeyama oracle (CPU only)."""
import numpy as np

from oracle import stats as S
from rmagine_amd import types as T


def test_paper_example():
    # tests/core/math_statistics.cpp:323-380: four 1-D correspondences, two merge levels
    c = [S.init_stats([1, 0, 0], [3, 0, 0]), S.init_stats([5, 0, 0], [1, 0, 0]),
         S.init_stats([6, 0, 0], [4, 0, 0]), S.init_stats([8, 0, 0], [8, 0, 0])]
    r = S.merge(S.merge(c[0], c[1]), S.merge(c[2], c[3]))
    assert r["n_meas"] == 4
    np.testing.assert_allclose(r["dataset_mean"], [5, 0, 0])
    np.testing.assert_allclose(r["model_mean"], [4, 0, 0])
    expect = np.zeros((3, 3))
    expect[0, 0] = 4.0
    np.testing.assert_allclose(r["covariance"], expect)
    # the direct two-pass formula gives the same
    d = S.statistics("p2p", T.identity_transforms(1), {"points": np.array([[1, 0, 0], [5, 0, 0], [6, 0, 0], [8, 0, 0]], np.float32)},
                     {"points": np.array([[3, 0, 0], [1, 0, 0], [4, 0, 0], [8, 0, 0]], np.float32)}, 4, 1, 1e9)[0]
    np.testing.assert_allclose(d["covariance"], expect)


def _reference_cloud(n=100000):
    # tests/core/math_statistics.cpp:497-509
    p = (np.arange(n, dtype=np.float64) / n).astype(np.float32)
    d = np.stack([-p, p * np.float32(10.0), p], 1).astype(np.float32)
    m = np.stack([p, p, -p], 1).astype(np.float32)
    mask = (np.arange(n) % 2).astype(np.uint8)
    ids = (np.arange(n) % 4).astype(np.uint32)
    normals = np.tile(np.array([0, 0, 1], np.float32), (n, 1))
    return d, m, mask, ids, normals


def test_filter_counts_p2p_p2l():
    # tests/core/math_statistics.cpp:479-660: n, n/2 (mask), 0 (mask & id 2)
    n = 100000
    d, m, mask, ids, normals = _reference_cloud(n)
    I = T.identity_transforms(1)
    for kind in ("p2p", "p2l"):
        md = {"points": m, "normals": normals}
        assert S.statistics(kind, I, {"points": d}, md, n, 1, 20000.0)[0]["n_meas"] == n
        assert S.statistics(kind, I, {"points": d, "mask": mask}, md, n, 1, 20000.0)[0]["n_meas"] == n // 2
        assert S.statistics(kind, I, {"points": d, "mask": mask, "ids": ids}, md, n, 1, 20000.0,
                            dataset_id=2)[0]["n_meas"] == 0


def test_merge_equals_direct_and_order():
    rng = np.random.default_rng(0)
    D = rng.normal(size=(500, 3))
    M = rng.normal(size=(500, 3)) + 3
    acc = S.identity_stats()
    for i in rng.permutation(500):
        acc = S.merge(acc, S.init_stats(D[i], M[i]))
    direct = S._direct(D, M)
    np.testing.assert_allclose(acc["covariance"], direct["covariance"], rtol=1e-10, atol=1e-12)
    np.testing.assert_allclose(acc["dataset_mean"], direct["dataset_mean"], rtol=1e-12)


def _random_rot(rng):
    q = rng.normal(size=4)
    q /= np.linalg.norm(q)
    return q


def test_umeyama_recovers_rigid_transform():
    rng = np.random.default_rng(1)
    for _ in range(20):
        q = _random_rot(rng)
        t = rng.uniform(-5, 5, 3)
        D = rng.uniform(-10, 10, size=(200, 3))
        M = D @ S.quat_matrix(q).T + t
        st = S._direct(D, M)
        r = S.umeyama(st)
        Rq = S.quat_matrix(r[:4])
        np.testing.assert_allclose(Rq, S.quat_matrix(q), atol=1e-12)
        np.testing.assert_allclose(r[4:], t, atol=1e-10)
        np.testing.assert_allclose(Rq.T @ Rq, np.eye(3), atol=1e-12)
        assert abs(np.linalg.det(Rq) - 1) < 1e-12


def test_umeyama_degenerate():
    assert np.array_equal(S.umeyama(S.identity_stats()), [0, 0, 0, 1, 0, 0, 0])
    # planar correspondences (rank 2) still give a proper rotation
    rng = np.random.default_rng(2)
    q = _random_rot(rng)
    D = np.c_[rng.uniform(-1, 1, (50, 2)), np.zeros(50)]
    M = D @ S.quat_matrix(q).T
    r = S.umeyama(S._direct(D, M))
    np.testing.assert_allclose(S.quat_matrix(r[:4]), S.quat_matrix(q), atol=1e-10)


def test_matrix_to_quat_branches():
    rng = np.random.default_rng(3)
    for _ in range(200):
        q = _random_rot(rng)
        q2 = S.matrix_to_quat(S.quat_matrix(q))
        assert min(np.abs(q2 - q).max(), np.abs(q2 + q).max()) < 1e-12
