"""Pins of the oracle against independent computations (CPU only).

* Pose algebra: tests/npref.py's FP32 restatement of Quaternion.tcc / Transform.tcc (which the
  oracle and the HIP kernels share, bit for bit) against scipy.spatial.transform.Rotation on
  random poses: rotation of vectors, composition, inversion.
* Sensor-frame attributes under rotated poses: oracle points and normals on a plane against the
  closed form built with scipy (normal = R^T e_z flipped toward the sensor, point = d t with
  t = h / -(R d)_z).
* Closest hit on >= 1e5 rays for each of four scene and sensor kinds (terrain + spherical,
  objects + pinhole, objects + random O1Dn, room with boxes + VLP-16): the oracle's BVH mode
  against its brute-force mode bit for bit, and against an FP64 Moller-Trumbore brute force
  (tests/cpp/mt_closest.c, a different algorithm) on every ray the oracle does not flag.
"""
import ctypes
import subprocess
from pathlib import Path

import numpy as np
import pytest

import oracle
from rmagine_amd import scenes
from rmagine_amd import types as T
from tests import npref

f4 = np.float32
MT_LIB = Path(__file__).resolve().parent / "cpp" / "_build" / "libmtcheck.so"


def _rng_poses(rng, n, trans_scale=5.0):
    q = rng.normal(size=(n, 4))
    q /= np.linalg.norm(q, axis=1, keepdims=True)
    return T.make_transforms(rng.uniform(-trans_scale, trans_scale, (n, 3)), q)


def test_pose_algebra_against_scipy():
    from scipy.spatial.transform import Rotation
    rng = np.random.default_rng(70)
    A, B = _rng_poses(rng, 300), _rng_poses(rng, 300)
    p = rng.normal(size=(300, 3)).astype(np.float32) * 10
    for i in range(300):
        qa = tuple(map(f4, A[i, :4]))
        Ra = Rotation.from_quat(A[i, :4].astype(np.float64))
        # rotate: q p q^-1
        np.testing.assert_allclose(npref.qrot(qa, p[i]), Ra.apply(p[i].astype(np.float64)), rtol=0, atol=1e-5)
        # compose: (A * B).R = Ra Rb, (A * B).t = Ra tb + ta
        C = npref.tmul(A[i], B[i])
        Rb = Rotation.from_quat(B[i, :4].astype(np.float64))
        ref = Ra * Rb
        q = C[:4].astype(np.float64)
        assert min(np.abs(q - ref.as_quat()).max(), np.abs(q + ref.as_quat()).max()) < 1e-6
        np.testing.assert_allclose(C[4:7], Ra.apply(B[i, 4:7].astype(np.float64)) + A[i, 4:7], atol=1e-5)
        # invert (Transform.tcc:24-31): R^-1 = conj, t^-1 = -(R^-1 t)
        qi = (-qa[0], -qa[1], -qa[2], qa[3])
        ti = -npref.qrot(qi, A[i, 4:7])
        np.testing.assert_allclose(ti, Ra.inv().apply(-A[i, 4:7].astype(np.float64)), atol=1e-5)
    # a sign flip in one cross term of the quaternion product would break the composition
    # above; the sandwich also equals scipy's rotation matrix applied to the unit axes
    q = tuple(map(f4, A[0, :4]))
    M = np.stack([npref.qrot(q, e) for e in np.eye(3, dtype=np.float32)], 1)
    np.testing.assert_allclose(M, Rotation.from_quat(A[0, :4].astype(np.float64)).as_matrix(), atol=1e-6)


def test_sensor_frame_points_and_normals_closed_form():
    """Oracle writers under rotated poses (SphereSimulatorEmbree.tcc:126-170): the point is the
    sensor-frame direction times the distance, the normal is the map normal rotated into the
    sensor frame and flipped toward the sensor; checked against scipy on a large plane."""
    from scipy.spatial.transform import Rotation
    V, F = scenes.genPlane()
    V = (np.asarray(V, np.float32) * f4(200.0)).astype(np.float32)  # z = 0, +-100 m
    om = oracle.OracleMap(V, F)
    rng = np.random.default_rng(71)
    n = 2000
    d = rng.normal(size=(n, 3))
    d = (d / np.linalg.norm(d, axis=1, keepdims=True)).astype(np.float32)
    model = T.O1DnModel(n, 1, T.Interval(0.0, 1000.0), (0.0, 0.0, 0.0), d)
    for k in range(8):
        P = _rng_poses(rng, 1, 3.0)
        P[0, 6] = f4(2.0 + k)  # height above the plane
        r = oracle.simulate(om, model, P)
        R = Rotation.from_quat(P[0, :4].astype(np.float64))
        dm = R.apply(d.astype(np.float64))
        t = (P[0, 6] / -dm[:, 2])
        hit = (t > 0) & np.all(np.abs(P[0, 4:6] + dm[:, :2] * t[:, None]) < 99.0, axis=1)
        ok = hit & (r["flags"] == 0)
        assert ok.sum() > 500
        np.testing.assert_array_equal(r["hits"][ok], 1)
        np.testing.assert_allclose(r["ranges"][ok], t[ok], rtol=2e-6)
        np.testing.assert_allclose(r["points"][ok], d[ok] * t[ok, None], rtol=0, atol=2e-6 * t[ok, None].max())
        ns = R.inv().apply([0.0, 0.0, 1.0])
        want = np.where((d[ok] @ ns)[:, None] > 0, -ns, ns)
        np.testing.assert_allclose(r["normals"][ok], want, atol=2e-6)


def _mt_lib():
    if not MT_LIB.exists():
        subprocess.run(["make", "-s", "-C", str(MT_LIB.parents[1]), "_build/libmtcheck.so"], check=True)
    lib = ctypes.CDLL(str(MT_LIB))
    P = ctypes.c_void_p
    lib.mt_closest.restype = None
    lib.mt_closest.argtypes = [P, P, ctypes.c_int64, P, P, ctypes.c_int64, ctypes.c_double, P, P]
    return lib


def _map_rays(model, poses):
    """FP32 map-frame ray origins and directions of the model from each pose (Tsb = identity),
    restated with npref's FP32 quaternion algebra, vectorised over rays."""
    if isinstance(model, T.SphericalModel):
        ds, os_ = npref.sphere_dirs(model), None
    elif isinstance(model, T.PinholeModel):
        ds, os_ = npref.pinhole_dirs(model), None
    else:
        ds, os_ = np.asarray(model.dirs, np.float32).reshape(-1, 3), np.asarray(model.orig, np.float32)
    O, D = [], []
    for P in poses:
        q = tuple(map(f4, P[:4]))
        D.append(npref.qrot_many(q, ds))
        o = P[4:7].astype(np.float32) if os_ is None else (npref.qrot(q, os_) + P[4:7]).astype(np.float32)
        O.append(np.broadcast_to(o, ds.shape))
    return np.concatenate(O).astype(np.float32), np.concatenate(D).astype(np.float32)


def _cases():
    rng = np.random.default_rng(72)
    Vt, Ft, Pt = scenes.terrain_workload(n_grid=160, n_sensors=1)
    Vo, Fo = scenes.objects_mesh(n_objects=300, n_ico=86, extent=40.0)
    q = rng.normal(size=(2, 4))
    q /= np.linalg.norm(q, axis=1, keepdims=True)
    Po = T.make_transforms(rng.uniform(-15, 15, (2, 3)), q)
    pin = T.PinholeModel(320, 240, T.Interval(0.1, 100.0), (160.0, 160.0), (159.5, 119.5))
    o1 = T.O1DnModel(100000, 1, T.Interval(0.1, 100.0), (0.0, 0.0, 0.0), scenes.random_unit_dirs(100000, rng))
    Vr, Fr, Pr = scenes.room_workload(n_sensors=7)
    return [("terrain spherical", Vt, Ft, scenes.ouster_workload_model(), Pt[:1]),
            ("objects pinhole", Vo, Fo, pin, Po),
            ("objects O1Dn", Vo, Fo, o1, Po[:1]),
            ("room VLP-16", Vr, Fr, scenes.vlp16_workload_model(), Pr)]


@pytest.fixture(scope="module")
def cases():
    out = []
    for name, V, F, model, P in _cases():
        om = oracle.OracleMap(V, F)
        r = oracle.simulate(om, model, P, attrs=("hits", "ranges", "face_ids"))
        out.append((name, V, F, model, P, om, r))
    return out


def test_bvh_equals_brute_force_1e5_rays(cases):
    for name, V, F, model, P, om, r in cases:
        assert len(r["ranges"]) >= 100000, name
        b = oracle.simulate(om, model, P, attrs=("hits", "ranges", "face_ids"), brute=True)
        for k in ("hits", "ranges", "face_ids", "flags"):
            np.testing.assert_array_equal(r[k], b[k], err_msg=f"{name} {k}")


def test_moller_trumbore_1e5_rays(cases):
    lib = _mt_lib()
    for name, V, F, model, P, om, r in cases:
        O, D = _map_rays(model, P)
        O64 = np.ascontiguousarray(O, np.float64)
        D64 = np.ascontiguousarray(D, np.float64)
        V32 = np.ascontiguousarray(V, np.float32)
        F32 = np.ascontiguousarray(F, np.uint32)
        n = len(O64)
        t = np.empty(n)
        f = np.empty(n, np.int64)
        lib.mt_closest(V32.ctypes.data, F32.ctypes.data, len(F32), O64.ctypes.data, D64.ctypes.data, n,
                       float(model.range.max), t.ctypes.data, f.ctypes.data)
        ok = r["flags"] == 0
        assert ok.mean() > 0.999, name
        hit_o = r["face_ids"] != T.UINT_MAX
        np.testing.assert_array_equal(hit_o[ok], (f >= 0)[ok], err_msg=name)
        both = ok & hit_o
        np.testing.assert_array_equal(r["face_ids"][both].astype(np.int64), f[both], err_msg=name)
        rel = np.abs(r["ranges"][both].astype(np.float64) - t[both]) / t[both]
        assert rel.max() < 1e-6, (name, rel.max())
        print(f"{name}: {n} rays, {int(hit_o.sum())} hits, {int((~ok).sum())} flagged, max rel range diff {rel.max():.2e}")
