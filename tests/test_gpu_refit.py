"""rmh_map_refit: after moving the vertices, simulation and closest points equal those of a
fresh build of the moved mesh bit for bit (the traversal is exact whatever the tree), and
the oracle on the moved mesh."""
import numpy as np
import pytest
import torch

import oracle
from rmagine_amd import scenes
from rmagine_amd import types as T
from rmagine_amd.simulation import HipMap, simulator_for

pytestmark = pytest.mark.gpu


def _sim(hm, model, P):
    r = simulator_for(model, hm).simulate(torch.as_tensor(P, device="cuda"), attrs=T.ATTRIBUTES)
    torch.cuda.synchronize()
    return r.to_numpy()


def _same(a, b, label):
    for k in T.ATTRIBUTES:
        np.testing.assert_array_equal(a[k].view(np.uint8), b[k].view(np.uint8), err_msg=f"{label}: {k}")


@pytest.mark.parametrize("motion", ["small", "large"])
def test_room_refit_equals_rebuild(motion):
    V, F, Tbm = scenes.room_workload(n_sensors=8)
    V = np.asarray(V, np.float32)
    rng = np.random.default_rng(1)
    if motion == "small":
        V2 = (V + rng.normal(scale=0.02, size=V.shape)).astype(np.float32)
    else:  # rotate the whole room by 30 degrees about z and shift it: every box moves far
        c, s = np.cos(0.5), np.sin(0.5)
        V2 = (V @ np.array([[c, -s, 0], [s, c, 0], [0, 0, 1]], np.float32).T + np.float32([1.5, -0.7, 0.2])).astype(np.float32)
    hm = HipMap(V, F)
    m = scenes.vlp16_workload_model()
    _sim(hm, m, Tbm)  # the tree was used before the update
    hm.refit(V2)
    got = _sim(hm, m, Tbm)
    _same(got, _sim(HipMap(V2, F), m, Tbm), f"room {motion}")
    o = oracle.simulate(oracle.OracleMap(V2, F), m, Tbm)
    for k in ("hits", "face_ids", "ranges"):
        np.testing.assert_array_equal(got[k], o[k], err_msg=f"oracle {k}")
    Q = torch.as_tensor(rng.uniform(-12, 12, size=(500, 3)).astype(np.float32), device="cuda")
    a, b = hm.closest_points(Q), HipMap(V2, F).closest_points(Q)
    for k in a:
        np.testing.assert_array_equal(a[k].cpu().numpy(), b[k].cpu().numpy(), err_msg=f"closest {k}")


def test_terrain_animation_repeated_refits():
    V, F, Tbm = scenes.terrain_workload(n_grid=96, n_sensors=4)
    V = np.asarray(V, np.float32)
    hm = HipMap(V, F)
    m = scenes.ouster_workload_model()
    for step in range(1, 4):
        Vs = V.copy()
        Vs[:, 2] = V[:, 2] * (1 + 0.3 * step) + 0.2 * np.sin(V[:, 0] * step)
        hm.refit(torch.as_tensor(Vs, device="cuda"))
        _same(_sim(hm, m, Tbm), _sim(HipMap(Vs, F), m, Tbm), f"terrain step {step}")


def test_refit_single_triangle_and_errors():
    V = np.array([[0, 0, 1], [1, 0, 1], [0, 1, 1]], np.float32)
    F = np.array([[0, 1, 2]], np.uint32)
    hm = HipMap(V, F)
    hm.refit(V + np.float32([0, 0, 1]))
    m = T.O1DnModel(1, 1, T.Interval(0.0, 10.0), (0.1, 0.1, 0.0), np.array([[0, 0, 1]], np.float32))
    P = np.array([[0, 0, 0, 1, 0, 0, 0, 0]], np.float32)
    r = _sim(hm, m, P)
    assert r["hits"][0] == 1 and abs(r["ranges"][0] - 2.0) < 1e-6
    with pytest.raises(ValueError):
        hm.refit(np.zeros((4, 3), np.float32))
