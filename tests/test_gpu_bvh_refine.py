"""SAH refinement of the radix tree (treelet restructuring, default on): the refined tree
has a lower SAH cost and every simulate() output equals the unrefined tree's bit for bit
(traversal only prunes; the FP64 triangle test decides); closest points agree too."""
import numpy as np
import pytest
import torch

from rmagine_amd import scenes
from rmagine_amd import types as T
from rmagine_amd.simulation import HipMap, simulator_for

pytestmark = pytest.mark.gpu


def _sim(hm, model, P):
    r = simulator_for(model, hm).simulate(torch.as_tensor(P, device="cuda"), attrs=T.ATTRIBUTES)
    torch.cuda.synchronize()
    return r.to_numpy()


@pytest.mark.parametrize("scene", ["room", "objects", "terrain"])
def test_refined_equals_radix_tree(scene):
    if scene == "room":
        V, F, P = scenes.room_workload(n_sensors=10)
        m = scenes.vlp16_workload_model()
    elif scene == "objects":
        V, F = scenes.objects_mesh(n_objects=600, n_ico=170, extent=60.0)
        P = T.make_transforms(np.random.default_rng(3).uniform(-20, 20, size=(6, 3)), [[0, 0, 0, 1]] * 6)
        m = scenes.vlp16_workload_model()
    else:
        V, F, P = scenes.terrain_workload(n_grid=192, n_sensors=3)
        m = scenes.ouster_workload_model()
    a, b = HipMap(V, F), HipMap(V, F, refine=False)
    ia, ib = a.info(), b.info()
    assert ia["sah_cost"] < ib["sah_cost"], (ia, ib)
    # both trees are stored compactly in preorder: the node array holds exactly the
    # reachable interior nodes (64 B each) beside 4 x 16 B of triangle data per face
    for i in (ia, ib):
        assert i["device_bytes"] == 64 * len(F) + 64 * i["n_nodes"], i
    ga, gb = _sim(a, m, P), _sim(b, m, P)
    for k in T.ATTRIBUTES:
        np.testing.assert_array_equal(ga[k].view(np.uint8), gb[k].view(np.uint8), err_msg=k)
    Q = torch.as_tensor(np.random.default_rng(1).uniform(np.min(V, 0), np.max(V, 0), size=(400, 3)).astype(np.float32),
                        device="cuda")
    ca, cb = a.closest_points(Q), b.closest_points(Q)
    for k in ca:
        np.testing.assert_array_equal(ca[k].cpu().numpy(), cb[k].cpu().numpy(), err_msg=f"closest {k}")
    print(scene, "sah", ib["sah_cost"], "->", ia["sah_cost"], "depth", ib["max_depth"], "->", ia["max_depth"])
