"""rmh_closest_points against the FP64 brute-force oracle (same FP32 map vertices)."""
import numpy as np
import pytest
import torch
from pathlib import Path

from oracle import closest as C
from rmagine_amd import scenes
from rmagine_amd.simulation import HipMap, import_hip_map, load_map_file

pytestmark = pytest.mark.gpu
MAPS = Path(__file__).resolve().parent / "golden" / "maps"


def _check(hm, V, F, Q, max_distance=C.FLT_MAX, label="", offsets=None):
    r = hm.closest_points(torch.as_tensor(Q, device="cuda"), max_distance)
    torch.cuda.synchronize()
    d, p, f = C.closest_points(V, F, Q, max_distance)
    f_glob = f.copy()
    if offsets is not None:  # the library reports the face index within its mesh (primID)
        off = np.asarray(offsets)
        f = np.where(f >= 0, f - off[np.searchsorted(off, f, side="right") - 1], f)
    gd = r["dist"].cpu().numpy()
    gf = r["face_ids"].cpu().numpy().astype(np.int64)
    np.testing.assert_array_equal(gd == C.FLT_MAX, d == C.FLT_MAX, err_msg=label)
    ok = d != C.FLT_MAX
    np.testing.assert_allclose(gd[ok], d[ok], rtol=2e-7, atol=1e-7, err_msg=label)
    np.testing.assert_allclose(r["points"].cpu().numpy()[ok], p[ok], atol=2e-6, err_msg=label)
    np.testing.assert_array_equal(gf[~ok], -1)
    # a different face than the oracle's is allowed only where the GPU's face is at the same
    # FP64 distance (a tie up to the rounding of the two FP64 evaluations, 1e-12 relative) and
    # is the lowest index among the tied faces
    V64 = np.asarray(V, np.float32).astype(np.float64)
    F64 = np.asarray(F, np.int64)
    if offsets is not None:  # compare in global face indices
        off = np.asarray(offsets)
        geom = r["geom_ids"].cpu().numpy().astype(np.int64)
        gf_glob = np.where(gf >= 0, gf + off[np.clip(geom, 0, len(off) - 1)], -1)
    else:
        gf_glob = gf
    diff = np.nonzero(ok & (gf_glob != f_glob))[0]
    for i in diff:
        P = C.closest_on_triangles(Q[i].astype(np.float64), V64[F64[:, 0]], V64[F64[:, 1]], V64[F64[:, 2]])
        d2 = ((P - Q[i].astype(np.float64)) ** 2).sum(1)
        tied = np.nonzero(d2 <= d2.min() * (1 + 1e-12) + 1e-300)[0]
        assert gf_glob[i] in tied, (label, i, gf_glob[i], f_glob[i], d2[gf_glob[i]], d2.min())
        assert gf_glob[i] == tied.min(), (label, i, gf_glob[i], tied)
    print(f"{label}: {ok.sum()} queries, {len(diff)} face ids differ from the oracle, all equal-distance ties")
    return r


def test_cube_known_answers_reference_scene():
    """The reference's closest-point test scene (two unit cubes at z = 5, x = 0 and 20)."""
    V, F = scenes.genCube()
    V = np.asarray(V, np.float32)
    V2 = np.concatenate([V + np.float32([0, 0, 5]), V + np.float32([20, 0, 5])])
    F2 = np.concatenate([F, F + 8]).astype(np.uint32)
    hm = HipMap(V2, F2, [0, 12, 24])
    Q = np.array([[0, 0, 0], [50, 50, 0], [50, -50, 20], [0.1, 0.1, 20], [-5, 0, 5]], np.float32)
    r = _check(hm, V2, F2, Q, label="cubes", offsets=[0, 12, 24])
    np.testing.assert_allclose(r["points"].cpu().numpy(),
                               [[0, 0, 4.5], [20.5, 0.5, 4.5], [20.5, -0.5, 5.5], [0.1, 0.1, 5.5], [-0.5, 0, 5]], atol=1e-6)
    np.testing.assert_array_equal(r["geom_ids"].cpu().numpy(), [0, 1, 1, 0, 0])


def test_room_random_queries_and_max_distance():
    V, F, Tbm = scenes.room_workload(n_sensors=1)
    hm = HipMap(V, F)
    rng = np.random.default_rng(9)
    Q = rng.uniform([-12, -12, -1], [12, 12, 5], size=(2000, 3)).astype(np.float32)
    _check(hm, V, F, Q, label="room")
    _check(hm, V, F, Q, max_distance=0.3, label="room max 0.3")
    r = hm.closest_points(torch.as_tensor(Q, device="cuda"), 0.0)
    assert (r["dist"].cpu().numpy() == C.FLT_MAX).all() and (r["face_ids"].cpu().numpy() == -1).all()


def test_imported_sphere_surface_and_empty():
    s = load_map_file(MAPS / "sphere.ply")
    hm = import_hip_map(MAPS / "sphere.ply")
    rng = np.random.default_rng(2)
    Q = rng.normal(size=(500, 3)).astype(np.float32) * 2
    _check(hm, s["vertices"], s["faces"], Q, label="sphere")
    empty = HipMap(np.zeros((0, 3), np.float32), np.zeros((0, 3), np.uint32))
    r = empty.closest_points(torch.as_tensor(Q[:5], device="cuda"))
    assert (r["dist"].cpu().numpy() == C.FLT_MAX).all()


def test_large_batch_in_morton_order_equals_caller_order():
    """Batches of 65,536 queries or more are walked in the Morton order of their cells within
    the map's box; every output equals that of the same queries in batches below that size,
    which are walked in the caller's order (queries far outside the box included)."""
    V, F, Tbm = scenes.room_workload(n_sensors=1)
    hm = HipMap(V, F)
    rng = np.random.default_rng(11)
    Q = rng.uniform([-14, -14, -3], [14, 14, 7], size=(150_000, 3)).astype(np.float32)
    Q[:100] *= 50
    Qd = torch.as_tensor(Q, device="cuda")
    for max_distance in (C.FLT_MAX, 0.3):
        big = hm.closest_points(Qd, max_distance)
        parts = [hm.closest_points(Qd[i:i + 50_000], max_distance) for i in range(0, len(Q), 50_000)]
        torch.cuda.synchronize()
        for k in big:
            np.testing.assert_array_equal(big[k].cpu().numpy(), torch.cat([p[k] for p in parts]).cpu().numpy(),
                                          err_msg=f"{k} max {max_distance}")
        sel = rng.choice(len(Q), 300, replace=False)
        d, p, f = C.closest_points(V, F, Q[sel], max_distance)
        gd = big["dist"].cpu().numpy()[sel]
        np.testing.assert_array_equal(gd == C.FLT_MAX, d == C.FLT_MAX)
        ok = d != C.FLT_MAX
        np.testing.assert_allclose(gd[ok], d[ok], rtol=2e-7, atol=1e-7)
        print(f"max {max_distance}: {len(Q)} queries equal in both orders, {ok.sum()} of 300 sampled hits match the oracle")


def test_large_batch_captured_into_a_graph():
    """Under stream capture a large batch is walked in the caller's order (no sort, no
    scratch): the replays of the captured call, with eager work between them, equal the
    eager call."""
    import ctypes
    from rmagine_amd import _lib, types as T
    from rmagine_amd.simulation import HipGraph, stream_handle
    V, F, Tbm = scenes.room_workload(n_sensors=1)
    hm = HipMap(V, F)
    rng = np.random.default_rng(12)
    Q = rng.uniform([-14, -14, -3], [14, 14, 7], size=(100_000, 3)).astype(np.float32)
    Qd = torch.as_tensor(Q, device="cuda")
    eager = hm.closest_points(Qd)
    names = ("dist", "points", "normals", "face_ids", "geom_ids", "object_ids")
    outs = {k: torch.empty_like(eager[k]) for k in names}
    out = T.rmh_closest_outputs(*(outs[k].data_ptr() for k in names))
    s = torch.cuda.Stream()
    g = HipGraph.capture(lambda st: _lib.call("rmh_closest_points", hm._handle, ctypes.c_void_p(Qd.data_ptr()),
                                              len(Q), ctypes.c_float(C.FLT_MAX), ctypes.byref(out),
                                              stream_handle(st)), s)
    for rep in range(3):
        for v in outs.values():
            v.zero_()
        torch.cuda.synchronize()
        g.launch(s)
        s.synchronize()
        for k in names:
            assert torch.equal(outs[k], eager[k]), (rep, k)
