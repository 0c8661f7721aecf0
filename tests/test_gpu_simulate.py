"""GPU parity of librmh's simulate() against the oracle (marked gpu).

Bar: hits, face/geom/object ids and ranges bit-exact (the kernel and the oracle evaluate
the same FP64 watertight predicate from the same FP32 rays); points and normals within
1e-5 (they are FP32 products of the same inputs and in practice bit-exact too).  Rays the
oracle flags as ambiguous for an FP32 intersector are counted and reported; they must
still match here.
"""
import numpy as np
import pytest
import torch

import oracle
from rmagine_amd import scenes
from rmagine_amd import types as T
from rmagine_amd.simulation import (Bundle, HipMap, O1DnSimulatorHip, OnDnSimulatorHip, PinholeSimulatorHip,
                                    SphereSimulatorHip, resize_memory_bundle, simulator_for)
from rmagine_amd._lib import NoMapError

pytestmark = pytest.mark.gpu

TOL = 1e-5  # m for ranges/points, rad-equivalent for unit normals


def run_gpu(map_, model, Tbm, attrs=T.ATTRIBUTES, Tsb=None):
    sim = simulator_for(model, map_)
    if Tsb is not None:
        sim.setTsb(Tsb)
    res = sim.simulate(torch.as_tensor(Tbm, device="cuda"), attrs=attrs)
    torch.cuda.synchronize()
    return res.to_numpy()


def compare(g, o, label=""):
    flags = o["flags"]
    report = {"rays": len(flags), "ambiguous": int((flags != 0).sum())}
    for k in ("hits", "face_ids", "geom_ids", "object_ids"):
        if k in g:
            bad = g[k] != o[k]
            report[k + "_mismatch"] = int(bad.sum())
            assert not bad.any(), f"{label}: {k} differs on {bad.sum()} rays ({(bad & (flags != 0)).sum()} flagged)"
    if "ranges" in g:
        d = np.abs(g["ranges"].astype(np.float64) - o["ranges"].astype(np.float64))
        report["range_max_err"] = float(d.max()) if len(d) else 0.0
        assert report["range_max_err"] <= TOL, f"{label}: range error {report['range_max_err']}"
    for k in ("points", "normals"):
        if k in g:
            np.testing.assert_array_equal(np.isnan(g[k]), np.isnan(o[k]), err_msg=f"{label}: {k} NaN pattern")
            ok = ~np.isnan(o[k])
            d = np.abs(g[k][ok].astype(np.float64) - o[k][ok]) if ok.any() else np.zeros(1)
            report[k + "_max_err"] = float(d.max())
            assert report[k + "_max_err"] <= TOL, f"{label}: {k} error {report[k + '_max_err']}"
    print(label, report)
    return report


@pytest.fixture(scope="module")
def cube():
    V, F = scenes.genCube()
    return V, F, HipMap(V, F), oracle.OracleMap(V, F)


def test_kat_spherical(cube):
    _, _, hm, om = cube
    m = T.example_spherical()
    Tbm = T.identity_transforms(100)
    g = run_gpu(hm, m, Tbm)
    r = g["ranges"][99 * m.size():][m.getBufferId(m.phi.size // 2, 0)]
    assert abs(r - 0.500076) < 1e-4  # tests/embree/simulation_spherical.cpp:57-66
    compare(g, oracle.simulate(om, m, Tbm), "cube/spherical")


def test_kat_pinhole_o1dn_ondn(cube):
    _, _, hm, om = cube
    Tbm = T.identity_transforms(100)
    for m in (T.example_pinhole(), T.example_o1dn(), T.example_ondn()):
        g = run_gpu(hm, m, Tbm)
        compare(g, oracle.simulate(om, m, Tbm), f"cube/{type(m).__name__}")
    g = run_gpu(hm, T.example_pinhole(), Tbm)
    assert np.abs(g["points"][0] - [0.5, 0.5, 0.375]).sum() < 1e-4   # simulation_pinhole.cpp:61-73
    g = run_gpu(hm, T.example_o1dn(), Tbm)
    assert np.abs(g["points"][0] - [0.256732, 0.5, -0.1]).sum() < 1e-4  # simulation_o1dn.cpp:59-72
    g = run_gpu(hm, T.example_ondn(), Tbm)
    assert abs(g["ranges"][0] - 0.5) < 1e-4  # simulation_ondn.cpp:63-70


@pytest.fixture(scope="module")
def room():
    V, F, Tbm = scenes.room_workload(n_sensors=100)
    return V, F, Tbm, HipMap(V, F), oracle.OracleMap(V, F)


def test_room_vlp16_all_attributes(room):
    """Room with 200 boxes (2,412 tris), 100 VLP-16 poses: 1.44M rays, every attribute."""
    V, F, Tbm, hm, om = room
    m = scenes.vlp16_workload_model()
    g = run_gpu(hm, m, Tbm)
    o = oracle.simulate(om, m, Tbm)
    rep = compare(g, o, "room/vlp16")
    assert rep["rays"] == 1_440_000
    assert g["hits"].mean() > 0.5


def test_room_other_models_and_tsb(room):
    V, F, Tbm, hm, om = room
    rng = np.random.default_rng(3)
    dirs = rng.normal(size=(3000, 3))
    dirs = (dirs / np.linalg.norm(dirs, axis=1, keepdims=True)).astype(np.float32)
    origs = rng.uniform(-0.2, 0.2, size=(3000, 3)).astype(np.float32)
    models = [
        T.PinholeModel(160, 120, T.Interval(0.05, 80.0), (80.0, 80.0), (79.5, 59.5)),
        T.O1DnModel(3000, 1, T.Interval(0.1, 60.0), (0.05, -0.02, 0.1), dirs),
        T.OnDnModel(1000, 3, T.Interval(0.1, 60.0), origs, dirs),
    ]
    Tsb = np.concatenate([T.quat_from_euler(0.02, -0.05, 0.3), [0.1, 0.0, 0.2, 0.0]]).astype(np.float32)
    for m in models:
        g = run_gpu(hm, m, Tbm[:20], Tsb=Tsb)
        o = oracle.simulate(om, m, Tbm[:20], Tsb=Tsb)
        compare(g, o, f"room/{type(m).__name__}")


def test_ray_list_order_independence(room):
    """O1Dn/OnDn rays are traced in direction-sorted order; results must follow each ray's
    own buffer index: permuting the ray list permutes every attribute the same way."""
    V, F, Tbm, hm, om = room
    rng = np.random.default_rng(11)
    n = 4099  # not a multiple of 64: a partial last tile
    dirs = rng.normal(size=(n, 3))
    dirs = (dirs / np.linalg.norm(dirs, axis=1, keepdims=True)).astype(np.float32)
    dirs[7] = 0.0          # zero direction: a miss, and a zero sort key
    dirs[8] = [0.0, 0.0, -1.0]
    origs = rng.uniform(-0.3, 0.3, size=(n, 3)).astype(np.float32)
    perm = rng.permutation(n)
    for make in (lambda d, o: T.O1DnModel(n, 1, T.Interval(0.1, 60.0), (0.05, -0.02, 0.1), d),
                 lambda d, o: T.OnDnModel(n, 1, T.Interval(0.1, 60.0), o, d)):
        a = run_gpu(hm, make(dirs, origs), Tbm[:6])
        b = run_gpu(hm, make(np.ascontiguousarray(dirs[perm]), np.ascontiguousarray(origs[perm])), Tbm[:6])
        for k in a:
            ra = a[k].reshape((6, n) + a[k].shape[1:])
            rb = b[k].reshape((6, n) + b[k].shape[1:])
            np.testing.assert_array_equal(rb, ra[:, perm], err_msg=k)
        o = oracle.simulate(om, make(dirs, origs), Tbm[:6])
        compare(a, o, "ray list")


def test_ray_order_switch(room):
    """The traversal order of O1Dn/OnDn rays (direction-sorted, or list order) changes no
    output byte; switching back and forth restores it; image models report list order."""
    V, F, Tbm, hm, om = room
    rng = np.random.default_rng(12)
    n = 2113
    dirs = rng.normal(size=(n, 3))
    dirs = (dirs / np.linalg.norm(dirs, axis=1, keepdims=True)).astype(np.float32)
    origs = rng.uniform(-0.3, 0.3, size=(n, 3)).astype(np.float32)
    for m in (T.O1DnModel(n, 1, T.Interval(0.1, 60.0), (0.05, -0.02, 0.1), dirs),
              T.OnDnModel(n, 1, T.Interval(0.1, 60.0), origs, dirs)):
        sim = simulator_for(m, hm)
        assert sim.ray_order_sorted()
        P = torch.as_tensor(Tbm[:5], device="cuda")
        ref = sim.simulate(P, attrs=T.ATTRIBUTES).to_numpy()
        for sorted_ in (False, True, False):
            sim.set_ray_order(sorted_)
            assert sim.ray_order_sorted() == sorted_
            got = sim.simulate(P, attrs=T.ATTRIBUTES).to_numpy()
            torch.cuda.synchronize()
            for k in ref:
                np.testing.assert_array_equal(got[k].view(np.uint8), ref[k].view(np.uint8),
                                              err_msg=f"{type(m).__name__} {k} sorted={sorted_}")
        compare(got, oracle.simulate(om, m, Tbm[:5]), f"list order/{type(m).__name__}")
    sim = simulator_for(T.PinholeModel(32, 24, T.Interval(0.05, 80.0), (20.0, 20.0), (15.5, 11.5)), hm)
    sim.set_ray_order(True)
    assert not sim.ray_order_sorted()


def test_walk_switch_and_sensor_tuner(room):
    """The walk of O1Dn/OnDn rays (speculative, or the classic while-while walk) changes no
    output byte in any ray order (direction-sorted, list, the W x H grid in 4 x 16 tiles), on
    scattered rays and on a scanner-like coherent grid whose sides 4 and 16 do not divide;
    autotune.tune_sensor applies one of the six pairs and changes nothing either; image
    models always walk the classic way; unknown orders are refused."""
    from rmagine_amd import autotune
    V, F, Tbm, hm, om = room
    rng = np.random.default_rng(13)
    n = 2113
    dirs = rng.normal(size=(n, 3))
    dirs = (dirs / np.linalg.norm(dirs, axis=1, keepdims=True)).astype(np.float32)
    origs = rng.uniform(-0.3, 0.3, size=(n, 3)).astype(np.float32)
    # 15 x 90 grid of a spinning scanner, row by row, beams starting off the centre
    ph, th = np.meshgrid(np.radians(np.linspace(-15.0, 15.0, 15)), np.radians(np.arange(90) * 4.0), indexing="ij")
    gd = np.stack([np.cos(ph) * np.cos(th), np.cos(ph) * np.sin(th), np.sin(ph)], -1).reshape(-1, 3).astype(np.float32)
    go = (0.015 * np.stack([np.cos(th), np.sin(th), np.zeros_like(th)], -1)).reshape(-1, 3).astype(np.float32)
    models = (T.O1DnModel(n, 1, T.Interval(0.1, 60.0), (0.05, -0.02, 0.1), dirs),
              T.OnDnModel(n, 1, T.Interval(0.1, 60.0), origs, dirs),
              T.O1DnModel(90, 15, T.Interval(0.1, 60.0), (0.0, 0.0, 0.0), gd),
              T.OnDnModel(90, 15, T.Interval(0.1, 60.0), go, gd))
    P = torch.as_tensor(Tbm[:5], device="cuda")
    for m in models:
        sim = simulator_for(m, hm)
        grid = m.getHeight() >= 4 and m.getWidth() >= 16  # a scanner grid: cast like an image
        assert sim.walk_speculative() == (not grid) and sim.ray_order() == ("tiles" if grid else "sorted")
        ref = sim.simulate(P, attrs=T.ATTRIBUTES).to_numpy()
        label = f"{type(m).__name__} {m.getWidth()}x{m.getHeight()}"
        for order in ("sorted", "list", "tiles"):
            for spec in (False, True):
                sim.set_ray_order(order)
                sim.set_walk(spec)
                assert sim.walk_speculative() == spec and sim.ray_order() == order
                got = sim.simulate(P, attrs=T.ATTRIBUTES).to_numpy()
                torch.cuda.synchronize()
                for k in ref:
                    np.testing.assert_array_equal(got[k].view(np.uint8), ref[k].view(np.uint8),
                                                  err_msg=f"{label} {k} order={order} speculative={spec}")
        compare(got, oracle.simulate(om, m, Tbm[:5]), f"walk/{label}")
        rec = autotune.tune_sensor(sim, P, reps=1)
        assert set(rec["trials"]) == {f"{o}/{w}" for o in ("sorted", "list", "tiles") for w in ("speculative", "classic")}
        assert sim.ray_order() == rec["ray_order"] and sim.walk_speculative() == rec["speculative"]
        assert rec["ms"] == min(rec["trials"].values()) or abs(rec["ms"] - min(rec["trials"].values())) < 1e-3
        got = sim.simulate(P, attrs=T.ATTRIBUTES).to_numpy()
        for k in ref:
            np.testing.assert_array_equal(got[k].view(np.uint8), ref[k].view(np.uint8), err_msg=f"{label} {k} tuned")
    with pytest.raises(ValueError):
        sim.set_ray_order("columns")
    from rmagine_amd import _lib
    assert _lib.lib().rmh_sensor_set_ray_order(sim._sensor.handle, 3) == 1  # RMH_ERR_INVALID
    sim = simulator_for(scenes.vlp16_workload_model(), hm)
    sim.set_walk(True)
    assert not sim.walk_speculative()
    sim.set_ray_order("tiles")
    assert sim.ray_order() == "list"
    assert autotune.tune_sensor(sim, P)["trials"] == {}


def test_attribute_subsets(room):
    V, F, Tbm, hm, om = room
    m = scenes.vlp16_workload_model()
    o = oracle.simulate(om, m, Tbm[:5])
    for attrs in (("ranges",), ("hits", "ranges"), ("normals", "face_ids"), ("points",)):
        g = run_gpu(hm, m, Tbm[:5], attrs=attrs)
        assert set(g) == set(attrs)
        compare(g, o, f"subset {attrs}")


def test_map_info_and_leaf_sizes(room):
    V, F, Tbm, _, om = room
    m = scenes.vlp16_workload_model()
    o = oracle.simulate(om, m, Tbm[:4])
    for leaf in (1, 2, 8):
        hm = HipMap(V, F, max_leaf_size=leaf)
        info = hm.info()
        assert info["n_faces"] == len(F) and 0 < info["max_depth"] <= 64
        compare(run_gpu(hm, m, Tbm[:4]), o, f"leaf{leaf}")


def test_terrain_small_parity():
    """Reduced terrain (128^2 grid, 32k tris), Ouster-style 128x1024 sensors, all attributes."""
    V, F, Tbm = scenes.terrain_workload(n_grid=128, n_sensors=8)
    hm = HipMap(V, F)
    om = oracle.OracleMap(V, F)
    m = scenes.ouster_workload_model()
    compare(run_gpu(hm, m, Tbm), oracle.simulate(om, m, Tbm), "terrain128")


def test_multi_mesh_geom_ids():
    V1, F1 = scenes.genCube()
    V2, F2 = scenes.genSphere(20, 40)
    V2 = V2 * np.float32(0.3)
    V = np.concatenate([V1, V2])
    F = np.concatenate([F1, F2 + len(V1)])
    off = [0, len(F1), len(F1) + len(F2)]
    hm, om = HipMap(V, F, off), oracle.OracleMap(V, F, off)
    m = T.example_spherical()
    Tbm = T.make_transforms([[0.35, 0.1, 0.0], [-0.4, -0.2, 0.1]], [[0, 0, 0, 1], [0, 0, 0.2, 0.98]])
    g = run_gpu(hm, m, Tbm)
    compare(g, oracle.simulate(om, m, Tbm), "multimesh")
    assert set(np.unique(g["geom_ids"])) == {0, 1}


def test_empty_map_and_degenerate():
    # a map without faces: every ray misses (the reference returns early, leaving outputs unwritten)
    hm = HipMap(np.zeros((0, 3), np.float32), np.zeros((0, 3), np.uint32))
    g = run_gpu(hm, T.example_spherical(), T.identity_transforms(2))
    assert (g["hits"] == 0).all() and (g["ranges"] == np.float32(101.0)).all() and (g["face_ids"] == T.UINT_MAX).all()
    # single triangle, and a degenerate (zero area) triangle next to it
    V = np.array([[1, -1, -1], [1, 1, -1], [1, 0, 1], [2, 0, 0], [2, 0, 0], [2, 1, 0]], np.float32)
    F = np.array([[0, 1, 2], [3, 4, 5]], np.uint32)
    hm1, om1 = HipMap(V[:3], F[:1]), oracle.OracleMap(V[:3], F[:1])
    hm2, om2 = HipMap(V, F), oracle.OracleMap(V, F)
    m = T.example_spherical()
    Tbm = T.identity_transforms(3)
    compare(run_gpu(hm1, m, Tbm), oracle.simulate(om1, m, Tbm), "single tri")
    compare(run_gpu(hm2, m, Tbm), oracle.simulate(om2, m, Tbm), "degenerate tri")


@pytest.mark.parametrize("wide", [False, True])
def test_exact_ties_on_shared_edges_and_vertices(wide):
    """Rays that land exactly on shared vertices, shared edges and cell diagonals of a
    triangle grid hit several triangles at the same distance.  Closed forms: every such
    ray hits (no crack between triangles) at t = 5, and the reported face is the lowest
    global face among the triangles whose closed region holds the point."""
    n = 6
    xs, ys = np.meshgrid(np.arange(n + 1), np.arange(n + 1), indexing="ij")
    V = np.stack([xs.ravel(), ys.ravel(), np.zeros(xs.size)], axis=1).astype(np.float32)
    vid = lambda i, j: i * (n + 1) + j
    F = []
    for i in range(n):
        for j in range(n):
            # alternate the diagonal so that vertices have 4 to 8 incident triangles
            if (i + j) % 2 == 0:
                F += [[vid(i, j), vid(i + 1, j), vid(i + 1, j + 1)], [vid(i, j), vid(i + 1, j + 1), vid(i, j + 1)]]
            else:
                F += [[vid(i, j), vid(i + 1, j), vid(i, j + 1)], [vid(i + 1, j), vid(i + 1, j + 1), vid(i, j + 1)]]
    F = np.array(F, dtype=np.uint32)
    rng = np.random.default_rng(4)
    F = F[rng.permutation(len(F))]  # face order unrelated to the Morton order of the build
    pts = []
    for i in range(n + 1):
        for j in range(n + 1):
            for dx, dy in ((0.0, 0.0), (0.5, 0.0), (0.0, 0.5), (0.5, 0.5)):
                if i + dx <= n and j + dy <= n:
                    pts.append((i + dx, j + dy))
    pts = np.array(pts, dtype=np.float64)

    def in_closed(p, tri):
        a, b, c = (V[k, :2].astype(np.float64) for k in tri)
        cr = lambda u, v, w: (v[0] - u[0]) * (w[1] - u[1]) - (v[1] - u[1]) * (w[0] - u[0])
        d = np.array([cr(a, b, p), cr(b, c, p), cr(c, a, p)])
        return (d >= 0).all() or (d <= 0).all()

    expect = np.array([min(k for k, tri in enumerate(F) if in_closed(p, tri)) for p in pts], dtype=np.uint32)
    hm = HipMap(V, F, wide=wide)
    om = oracle.OracleMap(V, F)
    for d in ((0.0, 0.0, -1.0), (1.0, 1.0, -1.0)):
        d = np.array(d, dtype=np.float32)
        origs = np.concatenate([pts, np.zeros((len(pts), 1))], axis=1).astype(np.float32) - 5 * d
        dirs = np.repeat(d[None], len(pts), axis=0)
        m = T.OnDnModel(len(pts), 1, T.Interval(0.0, 100.0), origs, dirs)
        g = run_gpu(hm, m, T.identity_transforms(1))
        assert g["hits"].all(), f"dir {d}: {int((g['hits'] == 0).sum())} rays fell through"
        np.testing.assert_array_equal(g["ranges"], np.float32(5.0), err_msg=f"dir {d}")
        np.testing.assert_array_equal(g["face_ids"], expect, err_msg=f"dir {d}: lowest face among tied hits")
        o = oracle.simulate(om, m, T.identity_transforms(1))
        np.testing.assert_array_equal(g["face_ids"], o["face_ids"])


def test_poses_split_over_launches(cube):
    """More (pose, tile) workgroups than one dispatch may hold (2^25): the poses are cast by
    several launches; poses on both sides of the split equal their single-pose results."""
    V, F, hm, om = cube
    rng = np.random.default_rng(21)
    d = rng.normal(size=(64, 3))
    model = T.O1DnModel(64, 1, T.Interval(0.01, 10.0), (0.0, 0.0, 0.0),
                        (d / np.linalg.norm(d, axis=1, keepdims=True)).astype(np.float32))
    n = (1 << 25) + 5  # one 64-ray tile per pose
    idx = torch.arange(n, device="cuda", dtype=torch.int64)
    Tbm = torch.zeros((n, 8), dtype=torch.float32, device="cuda")
    Tbm[:, 3] = 1.0
    Tbm[:, 4] = ((idx % 97) - 48).float() * 1e-3
    Tbm[:, 5] = ((idx % 89) - 44).float() * 1e-3
    sim = simulator_for(model, hm)
    res = sim.simulate(Tbm, attrs=("hits", "ranges"))
    torch.cuda.synchronize()
    for p in (0, (1 << 25) - 1, 1 << 25, n - 1):
        one = sim.simulate(Tbm[p:p + 1].contiguous(), attrs=("hits", "ranges"))
        torch.cuda.synchronize()
        np.testing.assert_array_equal(res.ranges[p * 64:(p + 1) * 64].cpu().numpy().view(np.uint32),
                                      one.ranges.cpu().numpy().view(np.uint32), err_msg=f"pose {p}")
        np.testing.assert_array_equal(res.hits[p * 64:(p + 1) * 64].cpu().numpy(), one.hits.cpu().numpy())
        o = oracle.simulate(om, model, Tbm[p:p + 1].cpu().numpy(), attrs=("hits", "ranges"))
        assert (o["hits"] == one.hits.cpu().numpy()).all() and one.hits.cpu().numpy().all()
    del res
    torch.cuda.empty_cache()


def test_zero_poses_and_errors(cube):
    _, _, hm, _ = cube
    sim = SphereSimulatorHip(hm)
    sim.setModel(T.example_spherical())
    res = sim.simulate(torch.zeros((0, 8), dtype=torch.float32, device="cuda"))
    assert res.ranges.numel() == 0
    nomap = SphereSimulatorHip()
    nomap.setModel(T.example_spherical())
    with pytest.raises(NoMapError):
        nomap.simulate(torch.as_tensor(T.identity_transforms(1), device="cuda"))
    with pytest.raises(ValueError):
        b = resize_memory_bundle(Bundle(), 10, 10, 1, ("ranges",))
        sim.simulate(torch.as_tensor(T.identity_transforms(1), device="cuda"), b)
    with pytest.raises(Exception):
        HipMap(np.zeros((3, 3), np.float32), np.array([[0, 1, 7]], np.uint32))  # index out of range


def test_repeat_is_deterministic(room):
    V, F, Tbm, hm, _ = room
    m = scenes.vlp16_workload_model()
    a = run_gpu(hm, m, Tbm[:10])
    b = run_gpu(hm, m, Tbm[:10])
    for k in a:
        np.testing.assert_array_equal(a[k].view(np.uint8), b[k].view(np.uint8))


def test_visit_stats_closed_forms(cube, room):
    """rmh_simulate_visit_stats: a one-triangle map is a single leaf (no node visits, one
    leaf and one triangle test per ray); on real trees the counts are consistent."""
    one = HipMap(np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0]], np.float32), np.array([[0, 1, 2]], np.uint32))
    sim = SphereSimulatorHip(one)
    m = T.example_spherical()
    sim.setModel(m)
    st = sim.visit_stats(T.identity_transforms(3))
    rays = 3 * m.size()
    assert st["rays"] == rays
    assert st["vector_node_visits"] == 0 and st["scalar_node_visits"] == 0 and st["wave_iterations"] == 0
    assert st["leaves"] == rays and st["triangle_tests"] == rays
    V, F, Tbm, hm, _ = room
    sim = SphereSimulatorHip(hm)
    sim.setModel(scenes.vlp16_workload_model())
    st = sim.visit_stats(Tbm[:10])
    nodes = st["vector_node_visits"] + st["scalar_node_visits"]
    assert nodes > st["rays"] and st["leaves"] >= 1 and st["triangle_tests"] >= st["leaves"]
    assert st["triangle_tests"] <= 8 * st["leaves"]
    assert 0 < st["uniform_wave_iterations"] <= st["wave_iterations"] <= nodes
    assert st["scalar_node_visits"] <= 64 * st["uniform_wave_iterations"]
    # a repeated call counts the same
    assert sim.visit_stats(Tbm[:10]) == st
