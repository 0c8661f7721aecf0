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


def _branch_rotations(rng):
    """Rotations that reach every branch of the matrix -> quaternion conversion: near the
    identity (trace > 0) and near half turns about x, y and z (the largest diagonal entry
    picks the branch)."""
    from scipy.spatial.transform import Rotation
    out = [Rotation.from_rotvec(rng.normal(size=3) * 0.3) for _ in range(20)]
    for axis in np.eye(3):
        for _ in range(20):
            v = axis * np.pi + rng.normal(size=3) * 0.05
            out.append(Rotation.from_rotvec(v))
    out += [Rotation.from_quat(q) for q in rng.normal(size=(200, 4))]
    return out


def test_pack_matrix_f32_against_scipy():
    """pack_matrix_f32 (the FP32 restatement of MatrixOps.tcc:1492-1529) against scipy's
    matrix -> quaternion: the same rotation up to sign, within FP32 rounding; every branch
    is reached."""
    rng = np.random.default_rng(10)
    rots = _branch_rotations(rng)
    M = np.stack([r.as_matrix() for r in rots]).astype(np.float32)
    q = S.pack_matrix_f32(M).astype(np.float64)
    ref = np.stack([r.as_quat() for r in rots])  # (x, y, z, w)
    err = np.minimum(np.abs(q - ref).max(1), np.abs(q + ref).max(1))
    assert err.max() < 2e-6, err.max()
    tr = np.trace(M, axis1=1, axis2=2)
    d = np.diagonal(M, axis1=1, axis2=2)
    branch = np.where(tr > 0, 0, np.where((d[:, 0] > d[:, 1]) & (d[:, 0] > d[:, 2]), 1, np.where(d[:, 1] > d[:, 2], 2, 3)))
    assert set(branch.tolist()) == {0, 1, 2, 3}
    # the largest quaternion component is the one the branch computes from the square root
    big = np.abs(q).argmax(1)
    np.testing.assert_array_equal(np.where(branch == 0, 3, branch - 1)[branch > 0], big[branch > 0])


def _objectwise_literal(bbox, width, n_points):
    """The reference kernel's loop over its 512 threads, literally
    (src/rmagine_cuda/src/math/statistics.cu:308-330), in 32-bit unsigned arithmetic."""
    min_col, min_row = S._f2u_sat(bbox[0]), S._f2u_sat(bbox[1])
    max_col, max_row = S._f2u_sat(bbox[3]), S._f2u_sat(bbox[4])
    bw, bh = (max_col - min_col) & 0xFFFFFFFF, (max_row - min_row) & 0xFFFFFFFF
    start = (min_row * width + min_col) & 0xFFFFFFFF
    n_elems = (bw * bh) & 0xFFFFFFFF
    per = (n_elems + 511) // 512
    ids = []
    for tid in range(512):
        for i in range(per):
            idx = tid * per + i
            row, col = idx // bw, idx % bw
            data_id = (start + row * width + col) & 0xFFFFFFFF
            if data_id < n_points:
                ids.append(data_id)
    return sorted(ids)


def test_objectwise_element_set_literal():
    width, height = 64, 40
    n = width * height
    boxes = [(3, 5, 0, 20, 9, 0),       # 68 elements: the tail runs 444 elements into later rows
             (0, 0, 0, 64, 8, 0),       # exactly 512
             (10, 30, 0, 50, 40, 0),    # reaches the scan's end: the tail is cut by n_points
             (5.9, 2.2, 0, 7.7, 3.1, 0),  # fractional corners truncate
             (-4, -1, 0, 2, 2, 0),      # negative corners saturate to 0
             (8, 8, 0, 8, 20, 0),       # zero width
             (0, 0, 0, 33, 17, 0)]      # 561 elements, rounded up to 1024
    for b in boxes:
        got = sorted(S.objectwise_elements(np.float32(b), width, n).tolist())
        assert got == _objectwise_literal(np.float32(b), width, n), b
    # window_only: exactly the window
    got = S.objectwise_elements(np.float32((3, 5, 0, 20, 9, 0)), width, n, window_only=True)
    grid = np.arange(n).reshape(height, width)
    assert sorted(got.tolist()) == sorted(grid[5:9, 3:20].ravel().tolist())


def test_objectwise_window_only_equals_plain_p2l():
    rng = np.random.default_rng(4)
    width, height = 50, 30
    n = width * height
    d = rng.normal(size=(n, 3)).astype(np.float32)
    m = (d + rng.normal(scale=0.05, size=(n, 3))).astype(np.float32)
    nr = rng.normal(size=(n, 3))
    nr = (nr / np.linalg.norm(nr, axis=1, keepdims=True)).astype(np.float32)
    mask = (rng.uniform(size=n) > 0.2).astype(np.uint8)
    ids = rng.integers(0, 3, n).astype(np.uint32)
    pre = T.make_transforms(rng.uniform(-0.1, 0.1, (2, 3)), [[0, 0, 0.05, 1], [0.02, 0, 0, 1]])
    boxes = np.float32([(4, 3, 0, 30, 20, 0), (0, 10, 0, 50, 30, 0)])
    params = [(0.5, 1, 1), (1e9, 2, 2)]
    ds = {"points": d, "mask": mask, "ids": ids}
    ms = {"points": m, "normals": nr, "ids": ids}
    got = S.statistics_objectwise_p2l(pre, ds, ms, width, n, params, boxes, window_only=True)
    grid = np.arange(n).reshape(height, width)
    for o, (b, (md, did, mid)) in enumerate(zip(boxes, params)):
        sel = grid[int(b[1]):int(b[4]), int(b[0]):int(b[3])].ravel()
        ref = S.statistics("p2l", pre[o], {"points": d[sel], "mask": mask[sel], "ids": ids[sel]},
                           {"points": m[sel], "normals": nr[sel], "ids": ids[sel]}, len(sel), 1, md, did, mid)[0]
        assert got[o]["n_meas"] == ref["n_meas"] > 0
        np.testing.assert_allclose(got[o]["covariance"], ref["covariance"], rtol=1e-12, atol=1e-14)
        np.testing.assert_allclose(got[o]["dataset_mean"], ref["dataset_mean"], rtol=1e-12)


def test_linear_system_finite_differences():
    """J^T w r is the gradient of (1/2) sum w r^2 (w held at its unperturbed values) and J^T w J
    is J^T W J of the residual Jacobian, both for the left perturbation Pd -> Exp(omega) Pd + tau
    of the pre-transformed dataset points; central differences in FP64."""
    from scipy.spatial.transform import Rotation
    rng = np.random.default_rng(7)
    n = 400
    d = rng.uniform(-5, 5, (n, 3)).astype(np.float32)
    nr = rng.normal(size=(n, 3))
    nr = (nr / np.linalg.norm(nr, axis=1, keepdims=True)).astype(np.float32)
    m = (d + rng.normal(scale=0.1, size=(n, 3))).astype(np.float32)
    pre = T.make_transforms([[0.1, -0.2, 0.05]], [[0.01, -0.02, 0.03, 1.0]])
    pre[0, :4] /= np.linalg.norm(pre[0, :4])
    for ks in (0.0, 5.0):
        sysm = S.linear_system_p2l(pre, {"points": d}, {"points": m, "normals": nr}, n, 1, 1e9, ks)[0]
        Pd0 = d.astype(np.float64) @ S.quat_matrix(pre[0, :4]).T + pre[0, 4:7].astype(np.float64)
        M64, N64 = m.astype(np.float64), nr.astype(np.float64)

        def residuals(xi):
            P = Rotation.from_rotvec(xi[:3]).apply(Pd0) + xi[3:]
            return np.einsum("ij,ij->i", P - M64, N64)

        r0 = residuals(np.zeros(6))
        w = np.ones(n) if ks <= 0 else ks ** 2 / (ks + r0) ** 2
        h = 1e-6
        Jfd = np.empty((n, 6))
        g = np.empty(6)
        for k in range(6):
            e = np.zeros(6)
            e[k] = h
            rp, rm = residuals(e), residuals(-e)
            Jfd[:, k] = (rp - rm) / (2 * h)
            g[k] = (0.5 * np.sum(w * rp * rp) - 0.5 * np.sum(w * rm * rm)) / (2 * h)
        np.testing.assert_allclose(sysm["JTwr"], g, rtol=1e-6, atol=1e-6)
        np.testing.assert_allclose(sysm["JTwJ"], (Jfd * w[:, None]).T @ Jfd, rtol=1e-6, atol=1e-5)
        assert sysm["n_meas"] == n
        np.testing.assert_allclose(sysm["wr2"], np.sum(w * r0 * r0), rtol=1e-12)
