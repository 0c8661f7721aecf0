"""GPU tests of the batched launch, graph replay, pose adjoint and batched RCC (marked gpu).

Bars: batched results bit-identical to per-sensor rmh_simulate (which is itself pinned to
the oracle) and to the oracle on sampled entries; graph replays bit-identical to eager
runs; adjoint gradients within 1e-9 (relative to the largest component) of an FP64 NumPy analytic
gradient built from the oracle's hits, and within 1e-3 of central finite differences
(h = 2^-14 m / 2^-12 rad) excluding rays whose hit face or Huber branch changes within +-h; batched gradients bit-identical to per-pose
launches; RCC converges like the reference test (|dz|, |yaw| <= 1e-3).
"""
import math

import numpy as np
import pytest
import torch

import oracle
from oracle import stats as S
from rmagine_amd import _lib
from rmagine_amd import adjoint as A
from rmagine_amd import icp
from rmagine_amd import math as M
from rmagine_amd import scenes
from rmagine_amd import types as T
from rmagine_amd.simulation import BatchSimulatorHip, HipGraph, HipMap, simulator_for

pytestmark = pytest.mark.gpu
dev = "cuda"


@pytest.fixture(scope="module")
def room():
    V, F, Tbm = scenes.room_workload(n_sensors=40)
    return V, F, Tbm, HipMap(V, F), oracle.OracleMap(V, F)


def _mixed_entries(Tbm, rng):
    models = [scenes.vlp16_workload_model(),
              T.PinholeModel(64, 48, T.Interval(0.05, 60.0), (32.0, 32.0), (31.5, 23.5))]
    for n in (1000, 4321, 64, 1):
        d = rng.normal(size=(n, 3))
        models.append(T.O1DnModel(n, 1, T.Interval(0.1, 60.0), (0.0, 0.0, 0.05),
                                  (d / np.linalg.norm(d, axis=1, keepdims=True)).astype(np.float32)))
    o = rng.uniform(-0.1, 0.1, (300, 3)).astype(np.float32)
    d = rng.normal(size=(300, 3))
    models.append(T.OnDnModel(100, 3, T.Interval(0.1, 60.0), o, (d / np.linalg.norm(d, axis=1, keepdims=True)).astype(np.float32)))
    entries = []
    for i, m in enumerate(models):
        k = [3, 5, 2, 7, 1, 4, 2][i]
        P = Tbm[i: i + k]
        entries.append({"model": m, "Tbm": torch.as_tensor(P, device=dev), "attrs": T.ATTRIBUTES,
                        "Tsb": np.concatenate([T.quat_from_euler(0, 0, 0.1 * i), [0.0, 0.01 * i, 0.1, 0.0]]).astype(np.float32)})
    return entries


def test_batch_matches_plain_and_oracle(room):
    V, F, Tbm, hm, om = room
    rng = np.random.default_rng(1)
    entries = _mixed_entries(Tbm, rng)
    batch = BatchSimulatorHip(hm, entries)
    for chunk, rows in ((256, 4), (64, 1), (1024, 16), (256, 64), (64, 8)):
        batch.set_params(chunk_rays=chunk, waves_per_cu=8 if chunk == 64 else 16, tile_rows=rows)
        assert batch.params()["tile_rows"] == rows
        bundles = batch.simulate()
        torch.cuda.synchronize()
        for e, b in zip(entries, bundles):
            sim = simulator_for(e["model"], hm)
            sim.setTsb(e["Tsb"])
            ref = sim.simulate(e["Tbm"], attrs=T.ATTRIBUTES)
            torch.cuda.synchronize()
            g, r = b.to_numpy(), ref.to_numpy()
            for k in T.ATTRIBUTES:
                np.testing.assert_array_equal(g[k].view(np.uint8), r[k].view(np.uint8), err_msg=f"{k} chunk {chunk} rows {rows}")
    for bad in ({"tile_rows": 3}, {"tile_rows": 128}, {"chunk_rays": 100}):
        with pytest.raises(_lib.RmhError):
            batch.set_params(**bad)
    assert batch.params()["tile_rows"] == 8  # a rejected call changes nothing
    # oracle on two entries
    for i in (1, 3):
        e = entries[i]
        o = oracle.simulate(om, e["model"], e["Tbm"].cpu().numpy(), Tsb=e["Tsb"])
        g = bundles[i].to_numpy()
        for k in ("hits", "face_ids", "ranges"):
            np.testing.assert_array_equal(g[k], o[k])


def test_batch_graph_replay(room):
    V, F, Tbm, hm, _ = room
    entries = _mixed_entries(Tbm, np.random.default_rng(2))
    batch = BatchSimulatorHip(hm, entries)
    s = torch.cuda.Stream()
    batch.simulate(stream=s)
    s.synchronize()
    eager = [b.to_numpy() for b in batch.bundles]
    for b in batch.bundles:
        for k in b.enabled():
            getattr(b, k).fill_(0)
    torch.cuda.synchronize()
    g = HipGraph.capture(lambda st: batch.simulate(stream=st), s)
    for rep in range(3):
        g.launch(s)
        s.synchronize()
        for e, b in zip(eager, batch.bundles):
            got = b.to_numpy()
            for k in e:
                np.testing.assert_array_equal(got[k].view(np.uint8), e[k].view(np.uint8), err_msg=f"replay {rep} {k}")
        for b in batch.bundles:  # eager work between replays
            b.ranges.fill_(-1.0)
        torch.cuda.synchronize()


def _target(sim, T_true, noise, seed):
    res = sim.simulate(torch.as_tensor(T_true, device=dev), attrs=("hits", "ranges"))
    torch.cuda.synchronize()
    r = res.ranges.clone()
    rng = np.random.default_rng(seed)
    r += torch.as_tensor(rng.normal(scale=noise, size=r.shape[0]).astype(np.float32), device=dev)
    mask = res.hits.clone()
    return r.contiguous(), mask.contiguous()


def _np_grad(o, model, Tbm, target, mask, V, F, delta, Tsb=None):
    """FP64 analytic gradient from the oracle's hits (independent of the kernel)."""
    Tsb = T.identity_transforms(1)[0] if Tsb is None else Tsb
    from tests import npref
    W = model.getWidth() * model.getHeight()
    dirs = npref.sphere_dirs(model)
    V64 = V.astype(np.float64)
    out = []
    for p in range(len(Tbm)):
        sl = slice(p * W, (p + 1) * W)
        Tsm = npref.tmul(Tbm[p], Tsb)
        q = tuple(map(np.float32, Tsm[0:4]))
        dm = np.stack([npref.qrot(q, d) for d in dirs]).astype(np.float64)
        om = Tsm[4:7].astype(np.float64)
        face = o["face_ids"][sl]
        hit = (o["hits"][sl] == 1) & (mask > 0)
        f = F[face[hit]].astype(np.int64)
        n = np.cross(V64[f[:, 1]] - V64[f[:, 0]], V64[f[:, 2]] - V64[f[:, 0]])
        n /= np.linalg.norm(n, axis=1, keepdims=True)
        n = n.astype(np.float32).astype(np.float64)
        d = dm[hit]
        # exact FP64 distance to the hit plane (the kernel keeps the FP64 distance too)
        ne = np.cross(V64[f[:, 1]] - V64[f[:, 0]], V64[f[:, 2]] - V64[f[:, 0]])
        r = np.einsum("ij,ij->i", ne, V64[f[:, 0]] - om) / np.einsum("ij,ij->i", ne, d)
        e = r - target[hit].astype(np.float64)
        ae = np.abs(e)
        rho = np.where(ae <= delta, 0.5 * e * e, delta * (ae - 0.5 * delta))
        drho = np.where(ae <= delta, e, delta * np.sign(e))
        nd = np.einsum("ij,ij->i", n, d)
        g = -n / nd[:, None]
        lever = om + r[:, None] * d - Tbm[p, 4:7].astype(np.float64)
        out.append(np.concatenate([[rho.sum()], (drho[:, None] * g).sum(0), (drho[:, None] * np.cross(lever, g)).sum(0),
                                   [hit.sum()]]))
    return np.asarray(out)


def test_pose_grad_vs_numpy_analytic(room):
    V, F, Tbm, hm, om = room
    model = scenes.vlp16_workload_model()
    sim = simulator_for(model, hm)
    target, mask = _target(sim, Tbm[:1], 0.01, 3)
    rng = np.random.default_rng(4)
    parts = np.repeat(Tbm[:1], 12, 0).copy()
    parts[:, 4:6] += rng.normal(scale=0.2, size=(12, 2)).astype(np.float32)
    for i in range(12):
        parts[i, 0:4] = npref_quat_mul(T.quat_from_euler(0, 0, rng.normal(scale=0.05)), parts[i, 0:4])
    g = A.pose_grad(sim, torch.as_tensor(parts, device=dev), target, mask, huber_delta=0.05).cpu().numpy()
    o = oracle.simulate(om, model, parts, attrs=("hits", "ranges", "face_ids"))
    # the threshold crosses the C ABI as an FP32 value
    ref = _np_grad(o, model, parts, target.cpu().numpy(), mask.cpu().numpy(), V, F, float(np.float32(0.05)))
    np.testing.assert_array_equal(g[:, 7], ref[:, 7])
    np.testing.assert_allclose(g[:, 0], ref[:, 0], rtol=1e-9)
    scale = np.abs(ref[:, 1:7]).max()
    err = np.abs(g[:, 1:7] - ref[:, 1:7]).max()
    print(f"pose grad vs FP64 NumPy analytic: max abs err {err:.3e} (scale {scale:.3e})")
    assert err <= 1e-9 * scale


def npref_quat_mul(a, b):
    from tests import npref
    return np.array(npref.qmul(tuple(map(np.float32, a)), tuple(map(np.float32, b))), np.float32)


def _pose_delta(Tp: np.ndarray, Tm: np.ndarray) -> np.ndarray:
    """(tau, omega) taking the float32 pose Tm to Tp: t_p - t_m and log(R_p R_m^T), FP64."""
    tp, tm = Tp[:, 4:7].astype(np.float64), Tm[:, 4:7].astype(np.float64)
    qp, qm = Tp[:, 0:4].astype(np.float64), Tm[:, 0:4].astype(np.float64)
    qp /= np.linalg.norm(qp, axis=1, keepdims=True)
    qm /= np.linalg.norm(qm, axis=1, keepdims=True)
    x1, y1, z1, w1 = qp.T
    x2, y2, z2, w2 = (-qm[:, 0], -qm[:, 1], -qm[:, 2], qm[:, 3])
    rel = np.stack([w1 * x2 + x1 * w2 + y1 * z2 - z1 * y2, w1 * y2 - x1 * z2 + y1 * w2 + z1 * x2,
                    w1 * z2 + x1 * y2 - y1 * x2 + z1 * w2, w1 * w2 - x1 * x2 - y1 * y2 - z1 * z2], 1)
    rel *= np.sign(rel[:, 3:4])
    v = rel[:, :3]
    nv = np.linalg.norm(v, axis=1, keepdims=True)
    ang = 2.0 * np.arctan2(nv, rel[:, 3:4])
    omega = np.where(nv > 0, v / np.where(nv > 0, nv, 1.0) * ang, 0.0)
    return np.concatenate([tp - tm, omega], 1)


def test_pose_grad_vs_finite_differences():
    """Central differences of the kernel's own loss, per component (h = 2^-14 m, 2^-12 rad).

    Rays whose hit face or Huber branch differs between -h, 0 and +h are excluded
    (through the per-pose target mask): the range field has kinks at triangle edges and the
    Huber loss is only C^1 at |e| = delta.  The measured difference
    L(+h) - L(-h) is compared with the analytic gradient dotted with the actual FP32 pose
    difference (translation difference and quaternion log), so that the rounding of the
    perturbed poses to FP32 does not enter the comparison.
    """
    V, F, _ = scenes.room_with_boxes(n_boxes=0)
    hm = HipMap(V, F)
    model = scenes.vlp16_workload_model()
    sim = simulator_for(model, hm)
    T_true = T.make_transforms([[1.0, -2.0, 1.2]], [T.quat_from_euler(0.02, -0.01, 0.4)])
    target, mask = _target(sim, T_true, 0.01, 5)
    rng = np.random.default_rng(6)
    n = 8
    T0 = np.repeat(T_true, n, 0).copy()
    T0[:, 4:7] += rng.normal(scale=0.2, size=(n, 3)).astype(np.float32)
    T0d = torch.as_tensor(T0, device=dev)
    tgt = target.repeat(n).contiguous()
    msk = mask.repeat(n)
    delta = float(np.float32(0.05))
    r0 = sim.simulate(T0d, attrs=("face_ids", "ranges"))
    f0, q0 = r0.face_ids, (r0.ranges - tgt).abs() <= delta
    for k in range(6):
        # rotations: the FP32 quaternions of the perturbed poses are unit only to ~1e-7 and
        # the sandwich product scales every direction by |q|^2 (an error ~1/h), while rays at
        # grazing incidence make the third-order term (~h^2) large; 2^-12 rad balances both
        h = 2.0 ** -14 if k < 3 else 2.0 ** -12
        xi = torch.zeros((n, 6), dtype=torch.float64, device=dev)
        xi[:, k] = h
        Tp, Tm = A.perturb(T0d, xi), A.perturb(T0d, -xi)
        rp = sim.simulate(Tp, attrs=("face_ids", "ranges"))
        rm = sim.simulate(Tm, attrs=("face_ids", "ranges"))
        # same hit face and same Huber branch at -h, 0, +h
        m = ((msk > 0) & (f0 == rp.face_ids) & (f0 == rm.face_ids) & (f0 != -1) &
             (q0 == ((rp.ranges - tgt).abs() <= delta)) & (q0 == ((rm.ranges - tgt).abs() <= delta)))
        m8 = m.to(torch.uint8).contiguous()
        an = A.pose_grad(sim, T0d, tgt, m8, per_pose_targets=True).cpu().numpy()[:, 1:7]
        lp = A.pose_grad(sim, Tp, tgt, m8, per_pose_targets=True).cpu().numpy()[:, 0]
        lm = A.pose_grad(sim, Tm, tgt, m8, per_pose_targets=True).cpu().numpy()[:, 0]
        dxi = _pose_delta(Tp.cpu().numpy(), Tm.cpu().numpy())
        fd = (lp - lm) / (2 * h)
        pred = np.einsum("ij,ij->i", an, dxi) / (2 * h)
        err = np.abs(fd - pred)
        print(f"component {k}: excluded {(~m).sum().item()} of {m.numel()} rays, "
              f"max |fd - pred| {err.max():.3e}, max |dL/dxi_k| {np.abs(an[:, k]).max():.3e}")
        assert err.max() <= 1e-3 * max(1.0, np.abs(an[:, k]).max())


def test_pose_grad_batched_equals_plain(room):
    V, F, Tbm, hm, _ = room
    model = scenes.vlp16_workload_model()
    sim = simulator_for(model, hm)
    target, mask = _target(sim, Tbm[:1], 0.01, 7)
    P = torch.as_tensor(Tbm[:16], device=dev)
    batched = A.pose_grad(sim, P, target, mask).cpu().numpy()
    plain = np.concatenate([A.pose_grad(sim, P[i:i + 1].contiguous(), target, mask).cpu().numpy() for i in range(16)])
    np.testing.assert_array_equal(batched, plain)


def test_autograd_function():
    """PoseRangeLoss: backward == J_l(omega)^T (kernel gradient at the perturbed pose), and a
    masked central difference through the Function itself at omega != 0."""
    V, F, _ = scenes.room_with_boxes(n_boxes=0)
    hm = HipMap(V, F)
    model = scenes.vlp16_workload_model()
    sim = simulator_for(model, hm)
    T_true = T.make_transforms([[0.5, 0.5, 1.0]], [T.quat_from_euler(0, 0, -0.3)])
    target, mask = _target(sim, T_true, 0.01, 8)
    n = 4
    T0 = torch.as_tensor(np.repeat(T_true, n, 0), device=dev)
    xi0 = torch.tensor([[0.1, -0.05, 0.02, 0.01, -0.02, 0.05],
                        [0.0, 0.0, 0.0, 0.0, 0.0, 0.3],
                        [-0.2, 0.1, 0.0, 0.03, 0.02, -0.1],
                        [0.0, 0.0, 0.0, 0.0, 0.0, 0.0]], dtype=torch.float64, device=dev)
    tgt = target.repeat(n).contiguous()
    msk = mask.repeat(n).contiguous()
    xi = xi0.clone().requires_grad_(True)
    loss = A.pose_range_loss(xi, T0, sim, tgt, msk, 0.05)
    (loss * torch.arange(1, n + 1, device=dev, dtype=torch.float64)).sum().backward()
    g = xi.grad.cpu().numpy()
    # 1. algebra: the kernel gradient at the perturbed pose, omega part through J_l^T
    raw = A.pose_grad(sim, A.perturb(T0, xi0), tgt, msk, per_pose_targets=True).cpu().numpy()
    J = A.so3_left_jacobian(xi0[:, 3:].cpu()).numpy()
    w = np.arange(1, n + 1)[:, None]
    np.testing.assert_allclose(g[:, :3], w * raw[:, 1:4], rtol=1e-12, atol=1e-12)
    np.testing.assert_allclose(g[:, 3:], w * np.einsum("nji,nj->ni", J, raw[:, 4:7]), rtol=1e-12, atol=1e-9)
    np.testing.assert_allclose(loss.detach().cpu().numpy(), raw[:, 0], rtol=0, atol=0)
    # 2. central differences through the Function on rays with a stable face / Huber branch
    delta = float(np.float32(0.05))
    for k in range(6):
        h = 2.0 ** -14 if k < 3 else 2.0 ** -12
        e = torch.zeros_like(xi0)
        e[:, k] = h
        P0, Pp, Pm = (A.perturb(T0, x) for x in (xi0, xi0 + e, xi0 - e))
        r = [sim.simulate(P, attrs=("face_ids", "ranges")) for P in (P0, Pp, Pm)]
        q = [(b.ranges - tgt).abs() <= delta for b in r]
        m = ((msk > 0) & (r[0].face_ids == r[1].face_ids) & (r[0].face_ids == r[2].face_ids) &
             (r[0].face_ids != -1) & (q[0] == q[1]) & (q[0] == q[2])).to(torch.uint8).contiguous()
        x = xi0.clone().requires_grad_(True)
        A.pose_range_loss(x, T0, sim, tgt, m, 0.05).sum().backward()
        with torch.no_grad():
            fd = ((A.pose_range_loss(xi0 + e, T0, sim, tgt, m, 0.05) -
                   A.pose_range_loss(xi0 - e, T0, sim, tgt, m, 0.05)) / (2 * h)).cpu().numpy()
        an = x.grad[:, k].cpu().numpy()
        print(f"autograd component {k}: max |fd - grad| {np.abs(fd - an).max():.3e}, max |grad| {np.abs(an).max():.3e}")
        assert np.abs(fd - an).max() <= 2e-3 * max(1.0, np.abs(an).max())


def test_rcc_converges_like_reference():
    # tests/embree/correction_rcc.cpp:68-180: sphere model 360x32 (-64..+60 deg) in the unit
    # cube, estimate perturbed by t.z = 0.1, yaw = 0.1; 10 x 5 iterations; |dz|, |yaw| <= 1e-3
    V, F = scenes.genCube()
    hm = HipMap(V, F)
    model = T.SphericalModel(phi=T.DiscreteInterval(T.f32(-64.0 * T.DEG_TO_RAD_F), T.f32(4.0 * T.DEG_TO_RAD_F), 32),
                             theta=T.DiscreteInterval(T.f32(-math.pi), T.f32(1.0 * T.DEG_TO_RAD_F), 360),
                             range=T.Interval(0.5, 130.0))
    sim = simulator_for(model, hm)
    data = sim.simulate(torch.as_tensor(T.identity_transforms(1), device=dev), attrs=("hits", "points"))
    torch.cuda.synchronize()
    est = T.make_transforms([[0.0, 0.0, 0.1]], [T.quat_from_euler(0.0, 0.0, 0.1)])
    rcc = icp.BatchedRCC(hm, model, data.points, data.hits, torch.as_tensor(est, device=dev), n_inner=5, max_dist=100.0)
    Tf = rcc.run(10).cpu().numpy()[0]
    q = Tf[0:4]
    yaw = math.atan2(2 * (q[3] * q[2] + q[0] * q[1]), 1 - 2 * (q[1] ** 2 + q[2] ** 2))
    assert abs(Tf[6]) <= 1e-3 and abs(yaw) <= 1e-3, (Tf, yaw)


def test_batched_rcc_graph_equals_eager_and_oracle(room):
    V, F, Tbm, hm, om = room
    rng = np.random.default_rng(9)
    d = rng.normal(size=(4096, 3))
    model = T.O1DnModel(4096, 1, T.Interval(0.1, 60.0), (0.0, 0.0, 0.0),
                        (d / np.linalg.norm(d, axis=1, keepdims=True)).astype(np.float32))
    sim = simulator_for(model, hm)
    T_true = Tbm[:1]
    data = sim.simulate(torch.as_tensor(T_true, device=dev), attrs=("hits", "points"))
    torch.cuda.synchronize()
    n = 32
    hyp = np.repeat(T_true, n, 0).copy()
    hyp[:, 4:6] += rng.normal(scale=0.1, size=(n, 2)).astype(np.float32)
    for i in range(n):
        hyp[i, 0:4] = npref_quat_mul(T.quat_from_euler(0, 0, rng.normal(scale=0.03)), hyp[i, 0:4])
    eager = icp.BatchedRCC(hm, model, data.points, data.hits, torch.as_tensor(hyp, device=dev))
    e1 = eager.run(2).cpu().numpy()
    graphed = icp.BatchedRCC(hm, model, data.points, data.hits, torch.as_tensor(hyp, device=dev))
    s = torch.cuda.Stream()
    graphed.capture(2, s)
    graphed.graphs[2].launch(s)
    s.synchronize()
    np.testing.assert_array_equal(graphed.Tbm.cpu().numpy(), e1)
    np.testing.assert_array_equal(graphed.grad.cpu().numpy(), eager.grad.cpu().numpy())
    # one graph per iteration count: 1 then 2 replayed iterations equal 3 eager ones, and a
    # replay waits for an update of Tbm queued on the caller's stream
    again = icp.BatchedRCC(hm, model, data.points, data.hits, torch.as_tensor(hyp, device=dev))
    again.run(1, graph=True)
    again.run(2, graph=True)
    assert sorted(again.graphs) == [1, 2]
    e3 = icp.BatchedRCC(hm, model, data.points, data.hits, torch.as_tensor(hyp, device=dev))
    np.testing.assert_array_equal(again.Tbm.cpu().numpy(), e3.run(3).cpu().numpy())
    again.Tbm.copy_(torch.as_tensor(hyp, device=dev))
    again.run(2, graph=True)
    np.testing.assert_array_equal(again.Tbm.cpu().numpy(), e1)
    # CPU oracle loop (oracle simulate, FP64 statistics and Umeyama), first two hypotheses
    from tests import npref
    dpts = data.points.cpu().numpy()
    dmask = data.hits.cpu().numpy()
    for hidx in range(2):
        Tcur = hyp[hidx].copy()
        for _ in range(2):
            o = oracle.simulate(om, model, Tcur[None], attrs=("hits", "points", "normals"))
            Tpre = T.identity_transforms(1)[0]
            for _ in range(5):
                st = S.statistics("p2l", Tpre, {"points": dpts, "mask": dmask},
                                  {"points": o["points"], "mask": o["hits"], "normals": o["normals"]}, 4096, 1, 100.0)[0]
                u = S.umeyama(st)
                Tnext = np.concatenate([u, [0.0]]).astype(np.float32)
                Tpre = npref.tmul(Tpre, Tnext)
            Tcur = npref.tmul(Tcur, Tpre)
        np.testing.assert_allclose(e1[hidx, :7], Tcur[:7], atol=2e-5)
    # the estimates moved toward the true pose
    err0 = np.linalg.norm(hyp[:, 4:6] - T_true[0, 4:6], axis=1)
    err1 = np.linalg.norm(e1[:, 4:6] - T_true[0, 4:6], axis=1)
    assert np.median(err1) < 0.5 * np.median(err0)
