"""Pins the CPU oracle (CPU only).

* the reference's known-answer rays on the unit cube
  (tests/embree/simulation_{spherical,o1dn,pinhole,ondn,single_ray}.cpp, tolerance 1e-4
  there; checked here against the closed forms as well);
* BVH mode == brute-force mode, bit for bit, on the room workload;
* closest hits agree with an independent float64 Moller-Trumbore brute force outside the
  flagged (edge / tie) rays;
* closed-form plane distances, miss values, FP32 ray set-up and multi-mesh ids.
"""
import math

import numpy as np
import pytest

import oracle
from rmagine_amd import scenes
from rmagine_amd import types as T
from tests import npref


@pytest.fixture(scope="module")
def cube():
    V, F = scenes.genCube()
    return V, F, oracle.OracleMap(V, F)


def test_kat_spherical(cube):
    # tests/embree/simulation_spherical.cpp:57-66: ray (phi.size/2, 0) -> 0.500076
    _, _, om = cube
    m = T.example_spherical()
    r = oracle.simulate(om, m, T.identity_transforms(100))
    rng = r["ranges"][99 * m.size():][m.getBufferId(m.phi.size // 2, 0)]
    assert abs(rng - 0.500076) < 1e-4
    # closed form: the ray leaves through x = -0.5, t = 0.5 / |d.x|
    d = npref.sphere_dirs(m)[m.getBufferId(8, 0)]
    assert rng == np.float32(0.5 / abs(float(d[0])))


def test_kat_o1dn(cube):
    # tests/embree/simulation_o1dn.cpp:59-72
    _, _, om = cube
    m = T.example_o1dn()
    r = oracle.simulate(om, m, T.identity_transforms(100))
    rng = r["ranges"][99 * m.size()]
    p = m.dirs[0].astype(np.float64) * rng
    assert np.abs(p - np.array([0.256732, 0.5, -0.1])).sum() < 1e-4
    np.testing.assert_array_equal(r["normals"][0], np.array([0, -1, 0], dtype=np.float32))


def test_kat_pinhole_edge(cube):
    # tests/embree/simulation_pinhole.cpp:61-73: pixel (0,0) hits a cube edge
    _, _, om = cube
    m = T.example_pinhole()
    r = oracle.simulate(om, m, T.identity_transforms(100))
    g = 99 * m.size()
    p = r["points"][g]
    assert np.abs(p - np.array([0.5, 0.5, 0.375])).sum() < 1e-4
    assert r["flags"][g] & oracle.FLAG_EDGE


def test_kat_ondn_and_single_ray(cube):
    # tests/embree/simulation_ondn.cpp:63-70, tests/optix/simulation_ondn.cpp:135,
    # tests/embree/simulation_single_ray.cpp:32-65
    _, _, om = cube
    m = T.example_ondn()
    r = oracle.simulate(om, m, T.identity_transforms(100))
    assert abs(r["ranges"][99 * m.size()] - 0.5) < 1e-4
    np.testing.assert_array_equal(r["normals"][0], np.array([-1, 0, 0], dtype=np.float32))
    single = T.OnDnModel(1, 1, T.Interval(0.0, 100.0), np.zeros((1, 3), np.float32),
                         np.array([[1.0, 0.0, 0.0]], np.float32))
    r = oracle.simulate(om, single, T.identity_transforms(100))
    assert np.all(r["ranges"] == 0.5)


def test_miss_values():
    V, F = scenes.genPlane()
    om = oracle.OracleMap(V, F)
    m = T.O1DnModel(2, 1, T.Interval(0.25, 7.0), (0.0, 0.0, 1.0),
                    np.array([[0, 0, 1], [0, 0, -1]], np.float32))
    r = oracle.simulate(om, m, T.identity_transforms(1))
    assert r["hits"][0] == 0 and r["ranges"][0] == np.float32(8.0)
    assert np.isnan(r["points"][0]).all() and np.isnan(r["normals"][0]).all()
    assert r["face_ids"][0] == T.UINT_MAX and r["geom_ids"][0] == T.UINT_MAX and r["object_ids"][0] == T.UINT_MAX
    assert r["hits"][1] == 1 and r["ranges"][1] == 1.0
    np.testing.assert_array_equal(r["points"][1], [0, 0, 0])
    np.testing.assert_array_equal(r["normals"][1], [0, 0, 1])


def test_range_min_keeps_range():
    # SphereSimulatorEmbree.tcc:110-124: hit only if t >= range.min, range written anyway
    V, F = scenes.genPlane()
    om = oracle.OracleMap(V, F)
    m = T.O1DnModel(1, 1, T.Interval(2.0, 7.0), (0.0, 0.0, 1.0), np.array([[0, 0, -1]], np.float32))
    r = oracle.simulate(om, m, T.identity_transforms(1))
    assert r["hits"][0] == 0 and r["ranges"][0] == 1.0 and r["face_ids"][0] != T.UINT_MAX


def test_plane_closed_form():
    V, F = scenes.genPlane()
    V = V * np.float32(100.0)
    om = oracle.OracleMap(V, F)
    rng = np.random.default_rng(5)
    n = 500
    d = rng.normal(size=(n, 3))
    d[:, 2] = -np.abs(d[:, 2]) - 0.2
    d = (d / np.linalg.norm(d, axis=1, keepdims=True)).astype(np.float32)
    h = np.float32(3.0)
    m = T.O1DnModel(n, 1, T.Interval(0.0, 1000.0), (0.0, 0.0, float(h)), d)
    r = oracle.simulate(om, m, T.identity_transforms(1))
    t_exact = (float(h) / -d[:, 2].astype(np.float64))
    inside = np.all(np.abs(d[:, :2] * t_exact[:, None]) < 49.0, axis=1)
    np.testing.assert_array_equal(r["hits"][inside], 1)
    assert np.all(np.abs(r["ranges"][inside] - t_exact[inside]) <= 1e-6 * t_exact[inside])
    assert np.all(r["ranges"][inside] == t_exact[inside].astype(np.float32))


@pytest.fixture(scope="module")
def room():
    V, F, Tbm = scenes.room_workload(n_sensors=6)
    return V, F, Tbm, oracle.OracleMap(V, F)


def test_bvh_equals_brute_force(room):
    V, F, Tbm, om = room
    m = scenes.vlp16_workload_model()
    a = oracle.simulate(om, m, Tbm[:3])
    b = oracle.simulate(om, m, Tbm[:3], brute=True)
    for k in ("hits", "ranges", "face_ids", "flags"):
        np.testing.assert_array_equal(a[k], b[k], err_msg=k)
    np.testing.assert_array_equal(np.isnan(a["points"]), np.isnan(b["points"]))


def test_against_moller_trumbore(room):
    V, F, Tbm, om = room
    m = scenes.vlp16_workload_model()
    dirs_s = npref.sphere_dirs(m)
    rng = np.random.default_rng(11)
    sel = rng.choice(len(dirs_s), 600, replace=False)
    r = oracle.simulate(om, m, Tbm[:1])
    T0 = npref.tmul(Tbm[0], T.identity_transforms(1)[0])
    q = tuple(map(np.float32, T0[0:4]))
    dm = np.stack([npref.qrot(q, dirs_s[i]) for i in sel])
    o = np.broadcast_to(T0[4:7], dm.shape)
    t_mt, f_mt = npref.closest_hits_mt(V, F, o, dm, float(m.range.max))
    fl = r["flags"][sel]
    ok = fl == 0
    assert ok.mean() > 0.95
    hit = f_mt >= 0
    np.testing.assert_array_equal((r["face_ids"][sel] != T.UINT_MAX)[ok], hit[ok])
    np.testing.assert_array_equal(r["face_ids"][sel][ok & hit], f_mt[ok & hit])
    rel = np.abs(r["ranges"][sel][ok & hit] - t_mt[ok & hit]) / t_mt[ok & hit]
    assert rel.max() < 1e-6


def test_ray_setup_matches_restatement(room):
    # points = d_s * t: recover d_s and compare with the FP32 restatement of sensor_models.h
    V, F, Tbm, om = room
    m = scenes.vlp16_workload_model()
    r = oracle.simulate(om, m, Tbm[:1])
    d = npref.sphere_dirs(m)
    hit = r["face_ids"] != T.UINT_MAX
    np.testing.assert_array_equal(r["points"][hit], (d[hit] * r["ranges"][hit, None]).astype(np.float32))
    pm = T.PinholeModel(64, 48, T.Interval(0.0, 100.0), (32.0, 32.0), (31.5, 23.5))
    rp = oracle.simulate(om, pm, Tbm[:1])
    dp = npref.pinhole_dirs(pm)
    hp = rp["face_ids"] != T.UINT_MAX
    np.testing.assert_array_equal(rp["points"][hp], (dp[hp] * rp["ranges"][hp, None]).astype(np.float32))


def test_multi_mesh_ids():
    V1, F1 = scenes.genCube()
    V2, F2 = scenes.genPlane()
    V2 = V2 + np.array([0, 0, -2], np.float32)
    V = np.concatenate([V1, V2])
    F = np.concatenate([F1, F2 + len(V1)])
    om = oracle.OracleMap(V, F, [0, len(F1), len(F1) + len(F2)])
    m = T.O1DnModel(2, 1, T.Interval(0.0, 10.0), (0.0, 0.0, 0.0), np.array([[0, 0, -1], [1, 0, 0]], np.float32))
    T0 = T.make_transforms([[0.0, 0.0, -1.0]], [[0, 0, 0, 1]])
    r = oracle.simulate(om, m, T0)
    assert r["geom_ids"][0] == 1 and r["object_ids"][0] == 1 and r["face_ids"][0] in (0, 1)
    assert r["ranges"][0] == 1.0


def test_diff_report_counts():
    """oracle/diff.py restates apps/rmagine_simulation_diff_calc/Main.cpp:83-147 (both hit,
    only one, neither, high differences, max and average over both-hit rays) and adds exact
    mismatch counts and the unflagged-mismatch count of the parity bar."""
    from oracle import diff as D
    nan = np.nan
    g = {"hits": np.array([1, 0, 1, 1, 0], np.uint8), "ranges": np.array([1, 101, 2, 3, 101], np.float32),
         "points": np.array([[0, 0, 0], [nan] * 3, [1, 1, 1], [2, 2, 2], [nan] * 3], np.float32),
         "face_ids": np.array([4, T.UINT_MAX, 2, 3, T.UINT_MAX], np.uint32)}
    o = {"hits": np.array([1, 1, 0, 1, 0], np.uint8), "ranges": np.array([1, 5, 101, 3.001, 101], np.float32),
         "points": np.array([[0, 0, 0], [1, 1, 1], [nan] * 3, [2, 2, 2.5], [nan] * 3], np.float32),
         "face_ids": np.array([4, 9, T.UINT_MAX, 3, T.UINT_MAX], np.uint32),
         "flags": np.array([0, oracle.FLAG_EDGE, oracle.FLAG_TIE | oracle.FLAG_RANGE, 0, 0], np.uint8)}
    r = D.compare(g, o, 100.0)
    assert r["diff"]["ranges"] == {"both_hit": 2, "only_a": 1, "only_b": 1, "neither": 1, "both_hit_high_diff": 1,
                                   "max_diff": pytest.approx(0.001, rel=1e-3),
                                   "avg_diff": pytest.approx(0.0005, rel=1e-3)}
    assert r["diff"]["points"]["both_hit_high_diff"] == 1 and r["diff"]["points"]["max_diff"] == pytest.approx(0.5)
    assert r["mismatch"] == {"hits": 2, "ranges": 3, "points": 3, "face_ids": 2}
    assert r["flagged"] == {"any": 2, "edge": 1, "near_miss": 0, "tie": 1, "range_limit": 1}
    assert r["mismatching_rays"] == 3 and r["mismatching_unflagged"] == 1
    m = D.merge(D.merge(None, r), r)
    assert m["rays"] == 10 and m["mismatching_unflagged"] == 2
    assert m["diff"]["ranges"]["both_hit"] == 4 and m["diff"]["ranges"]["avg_diff"] == pytest.approx(0.0005, rel=1e-3)
