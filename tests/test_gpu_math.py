"""GPU parity of librmh's math layer against the FP64 oracle (marked gpu).

Bars: pose algebra bit-exact against the FP32 restatement (same operation order, no FMA);
FP64 statistics/moments within 1e-12 relative of the two-pass NumPy oracle (absolute
1e-12 x data scale), filter counts exact; Umeyama within 1e-10 of the LAPACK-SVD oracle.
"""
import numpy as np
import pytest
import torch

from oracle import stats as S
from rmagine_amd import math as M
from rmagine_amd import types as T
from tests import npref

pytestmark = pytest.mark.gpu

dev = "cuda"


def rand_transforms(n, rng, scale=5.0):
    q = rng.normal(size=(n, 4))
    q /= np.linalg.norm(q, axis=1, keepdims=True)
    return T.make_transforms(rng.uniform(-scale, scale, (n, 3)), q)


def test_transform_algebra_bit_exact():
    rng = np.random.default_rng(0)
    A, B = rand_transforms(1000, rng), rand_transforms(1000, rng)
    Ad, Bd = torch.as_tensor(A, device=dev), torch.as_tensor(B, device=dev)
    C = M.transform_mul(Ad, Bd).cpu().numpy()
    ref = np.stack([npref.tmul(A[i], B[i]) for i in range(1000)])
    np.testing.assert_array_equal(C[:, :7], ref[:, :7])
    C1 = M.transform_mul(Ad, Bd[:1]).cpu().numpy()  # Nx1
    np.testing.assert_array_equal(C1[5, :7], npref.tmul(A[5], B[0])[:7])
    C2 = M.transform_mul(Ad[:1], Bd).cpu().numpy()  # 1xN
    np.testing.assert_array_equal(C2[7, :7], npref.tmul(A[0], B[7])[:7])
    Ai = M.transform_inv(Ad).cpu().numpy()
    for i in range(0, 1000, 97):
        q = tuple(map(np.float32, A[i, :4]))
        qi = (-q[0], -q[1], -q[2], q[3])
        np.testing.assert_array_equal(Ai[i, :4], np.array(qi, np.float32))
        np.testing.assert_array_equal(Ai[i, 4:7], -npref.qrot(qi, A[i, 4:7]))
    I = M.transform_mul(Ad, M.transform_inv(Ad)).cpu().numpy()
    np.testing.assert_allclose(I[:, :4], np.tile([0, 0, 0, 1], (1000, 1)), atol=1e-6)
    np.testing.assert_allclose(I[:, 4:7], 0, atol=1e-5)
    p = rng.normal(size=(1000, 3)).astype(np.float32)
    P = M.transform_points(Ad, torch.as_tensor(p, device=dev)).cpu().numpy()
    for i in range(0, 1000, 131):
        q = tuple(map(np.float32, A[i, :4]))
        np.testing.assert_array_equal(P[i], (npref.qrot(q, p[i]) + A[i, 4:7]).astype(np.float32))
    Q = Ad.clone()
    Q[:, :4] *= 3.0
    M.quat_normalize_(Q)
    np.testing.assert_allclose(Q[:, :4].cpu().numpy(), A[:, :4], atol=1e-6)


@pytest.mark.parametrize("seg,nseg", [(1_000_003, 1), (32768, 64), (100, 3000), (0, 2)])
def test_moments(seg, nseg):
    rng = np.random.default_rng(seg)
    a = rng.normal(size=(seg * nseg, 3)).astype(np.float32) * 10 + 50
    b = rng.normal(size=(seg * nseg, 3)).astype(np.float32)
    mask = (rng.uniform(size=seg * nseg) > 0.3).astype(np.uint8)
    ad, bd, md = (torch.as_tensor(x, device=dev) for x in (a, b, mask))
    if seg == 0:
        ad = torch.zeros((1, 3), device=dev)
        bd = torch.zeros((1, 3), device=dev)
        md = torch.zeros(1, dtype=torch.uint8, device=dev)
    r = M.vec3_moments(ad, bd, md, segment_size=seg, n_segments=nseg)
    o = S.moments(a, b, mask, seg, nseg) if seg else {"sum": np.zeros((nseg, 3)), "mean": np.zeros((nseg, 3)),
                                                       "cov": np.zeros((nseg, 9)), "count": np.zeros(nseg)}
    np.testing.assert_array_equal(r["count"].cpu().numpy(), o["count"])
    np.testing.assert_allclose(r["sum"].cpu().numpy(), o["sum"], rtol=1e-12, atol=1e-9)
    np.testing.assert_allclose(r["mean"].cpu().numpy(), o["mean"], rtol=1e-12, atol=1e-12)
    np.testing.assert_allclose(r["cov"].cpu().numpy(), o["cov"], rtol=1e-12, atol=1e-11)


def test_statistics_filter_counts():
    # tests/core/math_statistics.cpp:479-660 (n, n/2, 0) on the GPU, p2p and p2l
    n = 100000
    p = (np.arange(n, dtype=np.float64) / n).astype(np.float32)
    d = np.stack([-p, p * np.float32(10.0), p], 1).astype(np.float32)
    m = np.stack([p, p, -p], 1).astype(np.float32)
    mask = torch.as_tensor((np.arange(n) % 2).astype(np.uint8), device=dev)
    ids = torch.as_tensor((np.arange(n) % 4).astype(np.int32), device=dev)
    nrm = torch.as_tensor(np.tile(np.array([0, 0, 1], np.float32), (n, 1)), device=dev)
    dd, mm = torch.as_tensor(d, device=dev), torch.as_tensor(m, device=dev)
    I = torch.as_tensor(T.identity_transforms(1), device=dev)
    prm = M.UmeyamaReductionConstraints(max_dist=20000.0)
    for kind in (M.P2P, M.P2L):
        model = M.PointCloudView(mm, normals=nrm)
        s = M.statistics(kind, I, M.PointCloudView(dd), model, prm)
        assert M.stats_n_meas(s).item() == n
        s = M.statistics(kind, I, M.PointCloudView(dd, mask=mask), model, prm)
        assert M.stats_n_meas(s).item() == n // 2
        s = M.statistics(kind, I, M.PointCloudView(dd, mask=mask, ids=ids), model,
                         M.UmeyamaReductionConstraints(20000.0, 2, 0))
        assert M.stats_n_meas(s).item() == 0
    # paper example through the device path: cov(0,0) = 4
    D = torch.as_tensor(np.array([[1, 0, 0], [5, 0, 0], [6, 0, 0], [8, 0, 0]], np.float32), device=dev)
    Mp = torch.as_tensor(np.array([[3, 0, 0], [1, 0, 0], [4, 0, 0], [8, 0, 0]], np.float32), device=dev)
    st = M.stats_to_numpy(M.statistics(M.P2P, I, M.PointCloudView(D), M.PointCloudView(Mp),
                                       M.UmeyamaReductionConstraints(1e9)))
    assert st["n_meas"][0] == 4
    np.testing.assert_allclose(st["covariance"][0][0, 0], 4.0, rtol=1e-15)
    np.testing.assert_allclose(st["dataset_mean"][0], [5, 0, 0])
    np.testing.assert_allclose(st["model_mean"][0], [4, 0, 0])


@pytest.mark.parametrize("kind", ["p2p", "p2l"])
@pytest.mark.parametrize("seg,nseg,shared", [(32768, 40, True), (200_000, 1, False), (1000, 300, False)])
def test_statistics_segmented(kind, seg, nseg, shared):
    rng = np.random.default_rng(seg + nseg)
    pre = rand_transforms(nseg, rng, scale=0.5)
    nd = seg if shared else seg * nseg
    d = (rng.normal(size=(nd, 3)) * 5).astype(np.float32)
    mp = (rng.normal(size=(seg * nseg, 3)) * 5 + 0.1).astype(np.float32)
    nr = rng.normal(size=(seg * nseg, 3))
    nr = (nr / np.linalg.norm(nr, axis=1, keepdims=True)).astype(np.float32)
    dmask = (rng.uniform(size=nd) > 0.1).astype(np.uint8)
    mmask = (rng.uniform(size=seg * nseg) > 0.1).astype(np.uint8)
    max_dist = 6.0
    data = {"points": d, "mask": dmask, "stride": 0 if shared else None}
    model = {"points": mp, "mask": mmask, "normals": nr}
    o = S.statistics(kind, pre, data, model, seg, nseg, max_dist)
    T_ = lambda x: torch.as_tensor(x, device=dev)
    g = M.stats_to_numpy(M.statistics(M.P2P if kind == "p2p" else M.P2L, T_(pre),
                                      M.PointCloudView(T_(d), mask=T_(dmask), segment_stride=0 if shared else None),
                                      M.PointCloudView(T_(mp), mask=T_(mmask), normals=T_(nr)),
                                      M.UmeyamaReductionConstraints(max_dist), segment_size=seg, n_segments=nseg))
    n_o = np.array([x["n_meas"] for x in o])
    np.testing.assert_array_equal(g["n_meas"], n_o)
    for key in ("dataset_mean", "model_mean", "covariance"):
        ref = np.stack([x[key] for x in o])
        np.testing.assert_allclose(g[key], ref, rtol=1e-11, atol=1e-11 * 25)


def test_umeyama_against_oracle():
    rng = np.random.default_rng(5)
    n = 500
    stats = np.zeros((n, 16))
    ref = []
    for i in range(n):
        q = rng.normal(size=4)
        q /= np.linalg.norm(q)
        D = rng.uniform(-10, 10, size=(64, 3))
        Mp = D @ S.quat_matrix(q).T + rng.uniform(-3, 3, 3) + rng.normal(scale=0.05, size=(64, 3))
        st = S._direct(D, Mp)
        if i == 0:
            st = S.identity_stats()                      # n_meas = 0 -> identity
        if i == 1:
            st = dict(st, covariance=np.zeros((3, 3)))   # zero covariance -> identity rotation
        stats[i, 0:3], stats[i, 3:6] = st["dataset_mean"], st["model_mean"]
        stats[i, 6:15] = np.asarray(st["covariance"]).T.reshape(9)
        stats[i].view(np.uint32)[30] = st["n_meas"]
        ref.append(S.umeyama(st))
    ref = np.asarray(ref)
    sd = torch.as_tensor(stats, device=dev)
    out64 = torch.empty((n, 7), dtype=torch.float64, device=dev)
    Tf, _ = M.umeyama_transform(sd, out64=out64)
    g = out64.cpu().numpy()
    sign = np.sign(np.sum(g[:, :4] * ref[:, :4], axis=1, keepdims=True))
    np.testing.assert_allclose(g[:, :4] * sign, ref[:, :4], atol=1e-10)
    np.testing.assert_allclose(g[:, 4:], ref[:, 4:], atol=1e-9)
    np.testing.assert_array_equal(g[0], [0, 0, 0, 1, 0, 0, 0])
    np.testing.assert_array_equal(g[1, :4], [0, 0, 0, 1])
    np.testing.assert_allclose(Tf.cpu().numpy()[:, :7], g.astype(np.float32), atol=1e-6)
    print("umeyama max |dq|", np.abs(g[:, :4] * sign - ref[:, :4]).max(), "max |dt|", np.abs(g[:, 4:] - ref[:, 4:]).max())


def test_pack_bit_exact():
    """pack: matrix input bit-identical to the FP32 restatement of the reference's matrix ->
    quaternion steps (all four branches), quaternion input copied, stamps kept."""
    from tests.test_stats_oracle import _branch_rotations
    rng = np.random.default_rng(21)
    rots = _branch_rotations(rng)
    Rm = np.stack([r.as_matrix() for r in rots]).astype(np.float32)
    n = len(Rm)
    t = rng.normal(size=(n, 3)).astype(np.float32)
    out = torch.zeros((n, 8), dtype=torch.float32, device=dev)
    out.view(torch.int32)[:, 7] = torch.arange(n, dtype=torch.int32, device=dev) + 7
    M.pack(torch.as_tensor(Rm, device=dev), torch.as_tensor(t, device=dev), out=out)
    g = out.cpu().numpy()
    np.testing.assert_array_equal(g[:, :4], S.pack_matrix_f32(Rm))
    np.testing.assert_array_equal(g[:, 4:7], t)
    np.testing.assert_array_equal(g.view(np.int32)[:, 7], np.arange(n) + 7)
    q = rng.normal(size=(n, 4)).astype(np.float32)
    g2 = M.pack(torch.as_tensor(q, device=dev), torch.as_tensor(t, device=dev)).cpu().numpy()
    np.testing.assert_array_equal(g2[:, :4], q)
    np.testing.assert_array_equal(g2[:, 4:7], t)
    assert not g2[:, 7].any()


def _image_clouds(rng, width, height):
    n = width * height
    d = rng.normal(size=(n, 3)).astype(np.float32) * 3
    m = (d + rng.normal(scale=0.05, size=(n, 3))).astype(np.float32)
    nr = rng.normal(size=(n, 3))
    nr = (nr / np.linalg.norm(nr, axis=1, keepdims=True)).astype(np.float32)
    mask = (rng.uniform(size=n) > 0.15).astype(np.uint8)
    ids = rng.integers(0, 4, n).astype(np.uint32)
    return d, m, nr, mask, ids


@pytest.mark.parametrize("window_only", [False, True])
def test_statistics_objectwise_against_oracle(window_only):
    """statistics_objectwise_p2l against the oracle's element set (the reference kernel's,
    including its 512-rounded tail, or the window alone) and FP64 p2l statistics; corner
    cases: fractional, negative, NaN, inverted and empty windows, windows past the scan."""
    rng = np.random.default_rng(31)
    width, height = 320, 120
    n = width * height
    d, m, nr, mask, ids = _image_clouds(rng, width, height)
    boxes = np.float32([(10, 5, 0, 40, 25, 0), (0, 0, 0, 320, 120, 0), (300, 100, 0, 320, 120, 0),
                        (5.7, 3.2, 0, 17.9, 9.99, 0), (-5, -3, 0, 8, 4, 0), (np.nan, 2, 0, 9, 6, 0),
                        (50, 50, 0, 40, 60, 0), (7, 7, 0, 7, 50, 0), (100, 110, 0, 200, 400, 0),
                        (0, 119, 0, 320, 130, 0), (1, 2, 0, 2, 3, 0)])
    k = len(boxes)
    pre = rand_transforms(k, rng, scale=0.05)
    params = [(float(rng.uniform(0.05, 0.5)), int(rng.integers(0, 4)), int(rng.integers(0, 4))) for _ in range(k)]
    params[1] = (1e9, 0, 0)
    Td = lambda x: torch.as_tensor(x, device=dev)
    use_ids = [True, False]
    for with_ids in use_ids:
        ds = {"points": d, "mask": mask, "ids": ids if with_ids else None}
        ms = {"points": m, "normals": nr, "ids": ids if with_ids else None}
        o = S.statistics_objectwise_p2l(pre, ds, ms, width, n, params, boxes, window_only)
        g = M.stats_to_numpy(M.statistics_objectwise_p2l(
            Td(pre), M.PointCloudView(Td(d), mask=Td(mask), ids=Td(ids.astype(np.int32)) if with_ids else None),
            M.PointCloudView(Td(m), normals=Td(nr), ids=Td(ids.astype(np.int32)) if with_ids else None),
            width, height, [M.UmeyamaReductionConstraints(*p) for p in params], Td(boxes), window_only=window_only))
        np.testing.assert_array_equal(g["n_meas"], [x["n_meas"] for x in o])
        assert g["n_meas"][1] > (0.2 if with_ids else 0.8) * n
        for key in ("dataset_mean", "model_mean", "covariance"):
            ref = np.stack([x[key] for x in o])
            np.testing.assert_allclose(g[key], ref, rtol=1e-11, atol=1e-12)
    # the default element set really differs from the window on a window of 30 x 20 = 600
    e_ref = S.objectwise_elements(boxes[0], width, n)
    e_win = S.objectwise_elements(boxes[0], width, n, window_only=True)
    assert len(e_ref) == 1024 and len(e_win) == 600


@pytest.mark.parametrize("ks", [0.0, 5.0])
def test_linear_system_p2l_against_oracle(ks):
    rng = np.random.default_rng(41)
    seg, nseg = 20000, 12
    pre = rand_transforms(nseg, rng, scale=0.2)
    d = (rng.normal(size=(seg, 3)) * 4).astype(np.float32)
    mp = (rng.normal(size=(seg * nseg, 3)) * 4).astype(np.float32)
    nr = rng.normal(size=(seg * nseg, 3))
    nr = (nr / np.linalg.norm(nr, axis=1, keepdims=True)).astype(np.float32)
    dmask = (rng.uniform(size=seg) > 0.1).astype(np.uint8)
    mmask = (rng.uniform(size=seg * nseg) > 0.1).astype(np.uint8)
    md = 3.0
    o = S.linear_system_p2l(pre, {"points": d, "mask": dmask, "stride": 0}, {"points": mp, "mask": mmask, "normals": nr},
                            seg, nseg, md, ks)
    Td = lambda x: torch.as_tensor(x, device=dev)
    g = M.p2l_system_to_numpy(M.linear_system_p2l(
        Td(pre), M.PointCloudView(Td(d), mask=Td(dmask), segment_stride=0),
        M.PointCloudView(Td(mp), mask=Td(mmask), normals=Td(nr)), M.UmeyamaReductionConstraints(md),
        kernel_scale=ks, segment_size=seg, n_segments=nseg))
    np.testing.assert_array_equal(g["n_meas"], [x["n_meas"] for x in o])
    for key in ("JTwJ", "JTwr"):
        ref = np.stack([x[key] for x in o])
        scale = np.abs(ref).max()
        np.testing.assert_allclose(g[key], ref, rtol=1e-10, atol=1e-12 * scale)
    np.testing.assert_allclose(g["wr2"], [x["wr2"] for x in o], rtol=1e-10)
