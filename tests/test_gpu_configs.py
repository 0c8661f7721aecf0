"""Oracle parity of the HIP path on every benchmark workload at its own scene and sensor
shape (marked gpu).  The bars are those of the north star:

* hits, face/geom/object ids, ranges, points and normals bit-identical to the oracle
  (oracle/oracle.cpp, the CPU restatement of the Embree simulate loops) on every ray the
  oracle does not flag as ambiguous for an FP32 intersector (within 1e-6 of an edge, a near
  miss, a near tie, a range limit); every mismatching ray must be flagged.  The flag counts
  and the cross-backend report of apps/rmagine_simulation_diff_calc/Main.cpp:83-147
  (oracle/diff.py) are printed;
* pose gradients within 1e-9 (relative to the largest component) of the FP64 NumPy
  analytic gradient built from the oracle's hits, and consistent with central differences;
* batched RCC poses within 2e-5 of the CPU oracle loop (oracle simulate, FP64 statistics,
  LAPACK Umeyama), its Gauss-Newton gradient within 1e-9 of the FP64 oracle system.

Workloads: terrain 1024^2 (2,093,058 tris) with 8 of its 1,000 Ouster-style sensors, all
attributes; the 500,132-triangle objects scene with the full 2,000-sensor mixed batch in one
persistent launch, 36 of its entries (12 VLP-16, 12 pinhole 640x480, 12 O1Dn with >= 40k
rays) compared with all attributes; the 49,968-triangle room with 48 of its particles;
16 hypotheses x 32,768 O1Dn rays on the objects scene for 2 RCC iterations.
"""
import numpy as np
import pytest
import torch

import oracle
from oracle import diff as D
from oracle import stats as S
from rmagine_amd import adjoint as A
from rmagine_amd import icp
from rmagine_amd import math as M
from rmagine_amd import scenes
from rmagine_amd import types as T
from rmagine_amd.simulation import BatchSimulatorHip, HipMap, simulator_for
from tests import npref

pytestmark = pytest.mark.gpu
dev = "cuda"


def _check(name, g, o, range_max):
    r = D.compare(g, o, range_max)
    print(f"{name}: {r['rays']} rays, flagged {r['flagged']}, mismatching rays {r['mismatching_rays']} "
          f"(unflagged {r['mismatching_unflagged']}), mismatches {r['mismatch']}")
    for k, d in r["diff"].items():
        print(f"  {k}: {d}")
    assert r["mismatching_unflagged"] == 0, r
    return r


def test_terrain_full_size_all_attributes():
    """2,093,058-triangle terrain (tree depth above the 10 LDS stack rows), 8 poses spread
    over the 1,000 sensors, 128 x 1024 rays each, all seven attributes."""
    V, F, Tbm = scenes.terrain_workload()
    hm = HipMap(V, F)
    info = hm.info()
    assert info["n_faces"] == 2093058 and info["max_depth"] > 10
    model = scenes.ouster_workload_model()
    P = Tbm[[0, 1, 2, 3, 250, 500, 750, 999]]
    sim = simulator_for(model, hm)
    g = sim.simulate(torch.as_tensor(P, device=dev), attrs=T.ATTRIBUTES).to_numpy()
    o = oracle.simulate(oracle.OracleMap(V, F), model, P)
    r = _check("terrain", g, o, model.range.max)
    assert r["diff"]["hits"]["both_hit"] > 0.3 * r["rays"]  # the upper rows see the sky


def test_objects_mixed_batch_one_launch():
    """The full mixed batch (2,000 sensors, ~211M rays) in one persistent launch; 36 sampled
    entries carry all attributes and are compared with the oracle, 40 more by hits + ranges."""
    V, F, sensors = scenes.mixed_sensor_workload()
    kinds = np.array([m.kind for m, _ in sensors])
    sizes = np.array([m.size() for m, _ in sensors])
    rng = np.random.default_rng(12)
    vlp = rng.choice(np.nonzero(kinds == T.MODEL_SPHERICAL)[0], 12, replace=False)
    pin = rng.choice(np.nonzero(kinds == T.MODEL_PINHOLE)[0], 12, replace=False)
    o1 = rng.choice(np.nonzero((kinds == T.MODEL_O1DN) & (sizes >= 40000))[0], 12, replace=False)
    full = set(np.concatenate([vlp, pin, o1]).tolist())
    rest = rng.choice([i for i in range(len(sensors)) if i not in full], 40, replace=False)
    entries = [{"model": m, "Tbm": torch.as_tensor(P, device=dev),
                "attrs": T.ATTRIBUTES if i in full else ("hits", "ranges")} for i, (m, P) in enumerate(sensors)]
    hm = HipMap(V, F)
    batch = BatchSimulatorHip(hm, entries)
    assert batch.total_rays > 200_000_000
    batch.simulate()
    torch.cuda.synchronize()
    om = oracle.OracleMap(V, F)
    acc = None
    for i in sorted(full) + sorted(rest.tolist()):
        m, P = sensors[i]
        g = batch.bundles[i].to_numpy()
        o = oracle.simulate(om, m, P, attrs=tuple(g.keys()))
        r = D.compare(g, o, m.range.max)
        assert r["mismatching_unflagged"] == 0, (i, m.kind, r)
        acc = D.merge(acc, r)
    print(f"mixed batch, 76 entries: {acc['rays']} rays, flagged {acc['flagged']}, "
          f"mismatching rays {acc['mismatching_rays']}, mismatches {acc['mismatch']}")
    for k, d in acc["diff"].items():
        print(f"  {k}: {d}")
    assert acc["rays"] > 4_000_000


def test_fine_room_particles_pose_gradient():
    """49,968-triangle room, 48 of the particles: the batched pose gradient against the FP64
    NumPy analytic gradient from the oracle's hits, and against central differences of the
    kernel's own loss on 8 of them (rays whose face or Huber branch changes within +-h masked)."""
    from tests.test_gpu_batch_adjoint import _np_grad, _pose_delta, _target
    V, F, T_true, P = scenes.fine_room_workload()
    assert len(F) == 49968
    hm = HipMap(V, F)
    model = scenes.vlp16_workload_model()
    sim = simulator_for(model, hm)
    target, mask = _target(sim, T_true, 0.01, 3)
    parts = P[:48].copy()
    delta = float(np.float32(0.05))
    g = A.pose_grad(sim, torch.as_tensor(parts, device=dev), target, mask, huber_delta=0.05).cpu().numpy()
    o = oracle.simulate(oracle.OracleMap(V, F), model, parts, attrs=("hits", "ranges", "face_ids"))
    gs = sim.simulate(torch.as_tensor(parts, device=dev), attrs=("hits", "ranges", "face_ids")).to_numpy()
    _check("fine room particles", gs, o, model.range.max)
    ref = _np_grad(o, model, parts, target.cpu().numpy(), mask.cpu().numpy(), V, F, delta)
    np.testing.assert_array_equal(g[:, 7], ref[:, 7])
    np.testing.assert_allclose(g[:, 0], ref[:, 0], rtol=1e-9)
    scale = np.abs(ref[:, 1:7]).max()
    err = np.abs(g[:, 1:7] - ref[:, 1:7]).max()
    print(f"pose grad vs FP64 NumPy analytic, 48 particles: max abs err {err:.3e} (scale {scale:.3e})")
    assert err <= 1e-9 * scale
    # five-point central differences on 8 particles (the fourth-order stencil cancels the h^2
    # term).  A perturbed pose's FP32 quaternion is unit only to ~6e-8: the sandwich product
    # scales every direction by s = |q|^2 and so every range by 1/s, which shifts the loss by
    # dL/ds (s - 1) with dL/ds = -sum huber'(e) r; that shift, not a gradient error, is added to
    # the prediction (it is ~1e-3 of the rotation gradients here)
    n = 8
    T0d = torch.as_tensor(parts[:n], device=dev)
    W = model.size()
    tgt = target.repeat(n).contiguous()
    msk = mask.repeat(n)
    r0 = sim.simulate(T0d, attrs=("face_ids", "ranges"))
    f0, q0 = r0.face_ids, (r0.ranges - tgt).abs() <= delta
    for k in range(6):
        h = 2.0 ** -14 if k < 3 else 2.0 ** -12
        Tk = {}
        m = (msk > 0) & (f0 != -1)
        for s_ in (-2, -1, 1, 2):
            xi = torch.zeros((n, 6), dtype=torch.float64, device=dev)
            xi[:, k] = s_ * h
            Tk[s_] = A.perturb(T0d, xi)
            rs = sim.simulate(Tk[s_], attrs=("face_ids", "ranges"))
            # same hit face and same Huber branch at every stencil point
            m &= (f0 == rs.face_ids) & (q0 == ((rs.ranges - tgt).abs() <= delta))
        m8 = m.to(torch.uint8).contiguous()
        r64 = r0.ranges.double().view(n, W)
        e0 = r64 - tgt.double().view(n, W)
        dhub = torch.where(e0.abs() <= delta, e0, delta * torch.sign(e0))
        dL_ds = (-(dhub * r64) * m.view(n, W)).sum(1).cpu().numpy()
        an = A.pose_grad(sim, T0d, tgt, m8, per_pose_targets=True).cpu().numpy()[:, 1:7]
        L = {s_: A.pose_grad(sim, Tk[s_], tgt, m8, per_pose_targets=True).cpu().numpy()[:, 0] for s_ in Tk}
        fd = (8.0 * (L[1] - L[-1]) - (L[2] - L[-2])) / (12 * h)
        Tn = {s_: Tk[s_].cpu().numpy() for s_ in Tk}
        dxi = (8.0 * _pose_delta(Tn[1], Tn[-1]) - _pose_delta(Tn[2], Tn[-2])) / (12 * h)
        sq = {s_: (Tn[s_][:, 0:4].astype(np.float64) ** 2).sum(1) for s_ in Tn}
        pred = np.einsum("ij,ij->i", an, dxi) + dL_ds * (8.0 * (sq[1] - sq[-1]) - (sq[2] - sq[-2])) / (12 * h)
        e = np.abs(fd - pred)
        print(f"component {k}: masked {(~m).sum().item()} of {n * W} rays, max |fd - pred| {e.max():.3e}, "
              f"max |dL/dxi_k| {np.abs(an[:, k]).max():.3e}")
        assert e.max() <= 1e-3 * max(1.0, np.abs(an[:, k]).max())


def test_objects_batched_rcc_against_oracle_loop():
    """16 hypotheses x 32,768 random O1Dn rays on the 500,132-triangle scene, 2 RCC
    iterations (simulate, 5 x (p2l statistics, Umeyama, update), update, Gauss-Newton
    system), eager and graph-replayed, against the CPU oracle loop."""
    V, F = scenes.objects_mesh()
    hm = HipMap(V, F)
    rng = np.random.default_rng(4)
    model = T.O1DnModel(32768, 1, T.Interval(0.1, 100.0), (0.0, 0.0, 0.0), scenes.random_unit_dirs(32768, rng))
    sim = simulator_for(model, hm)
    T_true = T.make_transforms([[0.0, 0.0, 0.0]], [[0.0, 0.0, 0.0, 1.0]])
    data = sim.simulate(torch.as_tensor(T_true, device=dev), attrs=("hits", "points"))
    torch.cuda.synchronize()
    n = 16
    hyp = np.repeat(T_true, n, 0).copy()
    hyp[:, 4:7] += rng.normal(scale=0.2, size=(n, 3)).astype(np.float32)
    for i in range(n):
        a = 0.5 * rng.normal(scale=np.deg2rad(2.0))
        hyp[i, 0:4] = [0.0, 0.0, np.sin(a), np.cos(a)]
    max_dist = 0.5
    eager = icp.BatchedRCC(hm, model, data.points, data.hits, torch.as_tensor(hyp, device=dev), max_dist=max_dist)
    e1 = eager.run(2).cpu().numpy()
    graphed = icp.BatchedRCC(hm, model, data.points, data.hits, torch.as_tensor(hyp, device=dev), max_dist=max_dist)
    g1 = graphed.run(2, graph=True).cpu().numpy()
    np.testing.assert_array_equal(g1, e1)
    np.testing.assert_array_equal(graphed.grad.cpu().numpy(), eager.grad.cpu().numpy())
    # the oracle loop for every hypothesis
    om = oracle.OracleMap(V, F)
    dpts, dmask = data.points.cpu().numpy(), data.hits.cpu().numpy()
    for h in range(n):
        Tcur = hyp[h].copy()
        for _ in range(2):
            o = oracle.simulate(om, model, Tcur[None], attrs=("hits", "points", "normals"))
            Tpre = T.identity_transforms(1)[0]
            for _ in range(5):
                st = S.statistics("p2l", Tpre, {"points": dpts, "mask": dmask},
                                  {"points": o["points"], "mask": o["hits"], "normals": o["normals"]}, 32768, 1,
                                  float(np.float32(max_dist)))[0]
                Tpre = npref.tmul(Tpre, np.concatenate([S.umeyama(st), [0.0]]).astype(np.float32))
            Tcur = npref.tmul(Tcur, Tpre)
        np.testing.assert_allclose(e1[h, :7], Tcur[:7], atol=2e-5)
    # the Gauss-Newton system of the last iteration, at the device's own Tpre and model scan
    ms = eager.model_scan
    sys_ref = S.linear_system_p2l(eager.Tpre.cpu().numpy(), {"points": dpts, "mask": dmask, "stride": 0},
                                  {"points": ms.points.cpu().numpy(), "mask": ms.hits.cpu().numpy(),
                                   "normals": ms.normals.cpu().numpy()}, 32768, n, float(np.float32(max_dist)), 5.0)
    got = M.p2l_system_to_numpy(eager.grad)
    np.testing.assert_array_equal(got["n_meas"], [s["n_meas"] for s in sys_ref])
    for key in ("JTwJ", "JTwr"):
        ref = np.stack([s[key] for s in sys_ref])
        np.testing.assert_allclose(got[key], ref, rtol=1e-9, atol=1e-9 * np.abs(ref).max())
    err0 = np.linalg.norm(hyp[:, 4:7], axis=1)
    err1 = np.linalg.norm(e1[:, 4:7], axis=1)
    print(f"median translation error {np.median(err0):.4f} -> {np.median(err1):.6f} m; "
          f"|J^T w r| median {np.median(np.linalg.norm(got['JTwr'], axis=1)):.3e}")
    assert np.median(err1) < 0.5 * np.median(err0)
