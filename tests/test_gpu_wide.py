"""The 4-wide quantised BVH layout (RMH_BVH_WIDE / HipMap(wide=True) / set_layout).

Bar: with the ray kernels on the 4-wide tree, every output attribute is bit-identical to
the binary tree's (traversal only prunes; the FP64 triangle test decides) and hits, ids and
ranges equal the oracle's.  Scenes are chosen to stress the 8-bit grids: centimetre
triangles 2 km from the origin, where a grid origin needs all of the FP32 mantissa; a
perfectly flat floor (zero-thickness boxes, padded only by the binary tree's
outward rounding); a room and a terrain with deep trees.  The plain, batched and
pose-gradient kernels, a refit and layout switches are covered.
"""
import numpy as np
import pytest
import torch

import oracle
from rmagine_amd import adjoint as A
from rmagine_amd import scenes
from rmagine_amd import types as T
from rmagine_amd.simulation import BatchSimulatorHip, HipMap, simulator_for

pytestmark = pytest.mark.gpu
dev = "cuda"


def _sim(hm, model, P, attrs=T.ATTRIBUTES):
    r = simulator_for(model, hm).simulate(torch.as_tensor(P, device=dev), attrs=attrs)
    torch.cuda.synchronize()
    return r.to_numpy()


def _same(a, b, label):
    for k in a:
        np.testing.assert_array_equal(a[k].view(np.uint8), b[k].view(np.uint8), err_msg=f"{label}: {k}")


def _check(V, F, model, P, label, min_levels=2):
    wide, binary = HipMap(V, F, wide=True), HipMap(V, F)
    info = wide.info()
    assert wide.wide and info["n_wide_nodes"] > 0 and info["wide_levels"] >= min_levels, info
    assert not binary.wide and binary.info()["n_wide_nodes"] == 0
    g = _sim(wide, model, P)
    _same(g, _sim(binary, model, P), label)
    o = oracle.simulate(oracle.OracleMap(V, F), model, P)
    for k in ("hits", "face_ids", "geom_ids", "object_ids", "ranges"):
        np.testing.assert_array_equal(g[k], o[k], err_msg=f"{label}: oracle {k}")
    print(label, {k: info[k] for k in ("n_nodes", "max_depth", "n_wide_nodes", "wide_levels")},
          "hit rate", float(g["hits"].mean()))
    return g


def test_room_and_terrain():
    V, F, P = scenes.room_workload(n_sensors=12)
    _check(V, F, scenes.vlp16_workload_model(), P, "room", min_levels=5)
    V, F, P = scenes.terrain_workload(n_grid=128, n_sensors=4)
    _check(V, F, scenes.ouster_workload_model(), P, "terrain128", min_levels=5)


def test_far_from_origin_centimetre_triangles():
    """Room shrunk to 2 m x 2 m (boxes of 2-20 cm) and moved to (1500, -2000, 40) m: a
    2 cm node needs a 2^-13 m grid, whose origin must stay below 2^11 m in magnitude."""
    V, F, P = scenes.room_workload(n_sensors=6)
    off = np.array([1500.0, -2000.0, 40.0])
    Vf = (np.asarray(V, np.float64) * 0.1 + off).astype(np.float32)
    P = P.copy()
    P[:, 4:7] = (P[:, 4:7].astype(np.float64) * 0.1 + off).astype(np.float32)
    m = T.SphericalModel(phi=scenes.vlp16_workload_model().phi, theta=scenes.vlp16_workload_model().theta,
                         range=T.Interval(0.001, 20.0))
    g = _check(Vf, F, m, P, "far")
    assert g["hits"].mean() > 0.5


def test_flat_floor_grazing_rays():
    """A flat 64 x 64 grid of triangles at z = 0 (every box has zero thickness in z), cast
    from 1 mm and 1 m above it at shallow elevations."""
    n = 64
    xs = np.linspace(-3.2, 3.2, n + 1, dtype=np.float32)
    X, Y = np.meshgrid(xs, xs, indexing="ij")
    V = np.stack([X.ravel(), Y.ravel(), np.zeros(X.size, np.float32)], 1)
    idx = np.arange((n + 1) * (n + 1)).reshape(n + 1, n + 1)
    a, b, c, d = idx[:-1, :-1].ravel(), idx[1:, :-1].ravel(), idx[1:, 1:].ravel(), idx[:-1, 1:].ravel()
    F = np.concatenate([np.stack([a, b, c], 1), np.stack([a, c, d], 1)]).astype(np.uint32)
    m = T.SphericalModel(phi=T.DiscreteInterval(-0.5, 0.05, 11), theta=T.DiscreteInterval(-np.pi, 2 * np.pi / 360, 360),
                         range=T.Interval(0.0, 50.0))
    P = T.make_transforms([[0.1, -0.2, 0.001], [1.0, 1.0, 1.0], [-2.0, 0.5, 0.25]],
                          [T.quat_from_euler(0, 0, 0.3 * i) for i in range(3)])
    g = _check(V, F, m, P, "flat")
    assert g["hits"].any()


def test_batch_and_pose_grad_identical():
    V, F, Tbm = scenes.room_workload(n_sensors=24)
    wide, binary = HipMap(V, F, wide=True), HipMap(V, F)
    rng = np.random.default_rng(3)
    d = rng.normal(size=(777, 3))
    o1 = T.O1DnModel(777, 1, T.Interval(0.1, 60.0), (0.0, 0.0, 0.05),
                     (d / np.linalg.norm(d, axis=1, keepdims=True)).astype(np.float32))
    pin = T.PinholeModel(64, 48, T.Interval(0.05, 60.0), (32.0, 32.0), (31.5, 23.5))
    outs = []
    for hm in (wide, binary):
        entries = [{"model": mm, "Tbm": torch.as_tensor(Tbm[i * 6:(i + 1) * 6], device=dev), "attrs": T.ATTRIBUTES}
                   for i, mm in enumerate((scenes.vlp16_workload_model(), o1, pin))]
        bundles = BatchSimulatorHip(hm, entries).simulate()
        torch.cuda.synchronize()
        outs.append([b.to_numpy() for b in bundles])
    for i, (a, b) in enumerate(zip(*outs)):
        _same(a, b, f"batch entry {i}")
    model = scenes.vlp16_workload_model()
    sim_w, sim_b = simulator_for(model, wide), simulator_for(model, binary)
    res = sim_b.simulate(torch.as_tensor(Tbm[:1], device=dev), attrs=("hits", "ranges"))
    torch.cuda.synchronize()
    target = (res.ranges + 0.01).contiguous()
    P = torch.as_tensor(Tbm[:16], device=dev)
    gw = A.pose_grad(sim_w, P, target, res.hits).cpu().numpy()
    gb = A.pose_grad(sim_b, P, target, res.hits).cpu().numpy()
    np.testing.assert_array_equal(gw, gb)
    assert np.abs(gw).max() > 0


def test_refit_and_layout_switch():
    V, F, Tbm = scenes.room_workload(n_sensors=6)
    V = np.asarray(V, np.float32)
    m = scenes.vlp16_workload_model()
    hm = HipMap(V, F)
    base = _sim(hm, m, Tbm)
    assert hm.set_layout(True) and hm.wide
    _same(_sim(hm, m, Tbm), base, "switched to wide")
    rng = np.random.default_rng(5)
    V2 = (V + rng.normal(scale=0.05, size=V.shape)).astype(np.float32)
    hm.refit(V2)
    assert hm.wide
    got = _sim(hm, m, Tbm)
    _same(got, _sim(HipMap(V2, F, wide=True), m, Tbm), "refit vs fresh wide build")
    o = oracle.simulate(oracle.OracleMap(V2, F), m, Tbm)
    for k in ("hits", "face_ids", "ranges"):
        np.testing.assert_array_equal(got[k], o[k], err_msg=f"refit oracle {k}")
    assert not hm.set_layout(False) and not hm.wide
    _same(_sim(hm, m, Tbm), got, "switched back to binary")
    # a single triangle is a single leaf: no wide tree, the layout stays binary
    one = HipMap(np.eye(3, dtype=np.float32), np.array([[0, 1, 2]], np.uint32), wide=True)
    assert not one.wide
    assert not one.set_layout(True)
