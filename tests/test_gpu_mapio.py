"""import_hip_map: simulation on imported maps against the oracle, ids included.

The oracle gets the same flattened arrays and per-mesh ids; hits, ranges, face/geom/object
ids must agree bit for bit (the map files are the reference's own dat/ files, copied under
tests/golden/maps, plus a generated COLLADA scene with a multi-mesh instance).
"""
from pathlib import Path

import numpy as np
import pytest
import torch

import oracle
from rmagine_amd import scenes
from rmagine_amd.ouster import o1dn_from_ouster_meta_file
from rmagine_amd import types as T
from rmagine_amd.simulation import import_hip_map, load_map_file, simulator_for

pytestmark = pytest.mark.gpu

MAPS = Path(__file__).resolve().parent / "golden" / "maps"


def _both(path, polygons="first"):
    s = load_map_file(path, polygons)
    om = oracle.OracleMap(s["vertices"], s["faces"], s["mesh_face_offsets"], s["geom_ids"], s["object_ids"])
    return import_hip_map(path, polygons), om, s


def _poses(n, lo, hi, seed):
    rng = np.random.default_rng(seed)
    P = np.zeros((n, 8), np.float32)  # rmh_transform: quaternion (x, y, z, w), t, stamp
    yaw = rng.uniform(-np.pi, np.pi, n)
    P[:, 2] = np.sin(yaw / 2)
    P[:, 3] = np.cos(yaw / 2)
    P[:, 4:7] = rng.uniform(lo, hi, size=(n, 3))
    return P


def _compare(hm, om, model, P, label):
    sim = simulator_for(model, hm)
    g = sim.simulate(torch.as_tensor(P, device="cuda"), attrs=T.ATTRIBUTES)
    torch.cuda.synchronize()
    g = g.to_numpy()
    o = oracle.simulate(om, model, P)
    for k in ("hits", "face_ids", "geom_ids", "object_ids"):
        np.testing.assert_array_equal(g[k], o[k], err_msg=f"{label}: {k}")
    np.testing.assert_array_equal(g["ranges"], o["ranges"], err_msg=f"{label}: ranges")
    return g


def test_sphere_ply_from_inside():
    hm, om, s = _both(MAPS / "sphere.ply")
    P = _poses(6, -0.3, 0.3, 1)
    g = _compare(hm, om, scenes.vlp16_workload_model(), P, "sphere.ply")
    assert g["hits"].all()
    # poses within 0.3*sqrt(3) of the centre of a unit sphere whose faces are > 0.98 away
    assert g["ranges"].min() > 0.98 - 0.53 and g["ranges"].max() < 1.0 + 0.53


@pytest.mark.parametrize("name", ["two_cubes.dae", "many_objects.dae", "box_rot_trans_scaled.dae"])
def test_collada_scenes(name):
    hm, om, s = _both(MAPS / name)
    V = s["vertices"]
    c = 0.5 * (V.min(0) + V.max(0))
    P = _poses(8, c - 4.0, c + 4.0, 2)
    g = _compare(hm, om, scenes.vlp16_workload_model(), P, name)
    if name == "many_objects.dae":
        assert len(np.unique(g["geom_ids"][g["hits"] == 1])) > 2


def test_multi_mesh_instance_ids(tmp_path):
    """A node with two meshes stays an instance: geom id = index inside it, object id =
    its instance id; the single-mesh node before it is frozen with its own id."""
    def geom(gid, z):
        return (f'<geometry id="{gid}"><mesh><source id="{gid}-p"><float_array id="{gid}-a" count="9">'
                f'-5 -5 {z} 5 -5 {z} 0 5 {z}</float_array></source><vertices id="{gid}-v">'
                f'<input semantic="POSITION" source="#{gid}-p"/></vertices><triangles count="1">'
                f'<input semantic="VERTEX" source="#{gid}-v" offset="0"/><p>0 1 2</p></triangles></mesh></geometry>')
    dae = ('<?xml version="1.0"?><COLLADA><library_geometries>' + geom("G0", 1.0) + geom("G1", 2.0) + geom("G2", 3.0) +
           '</library_geometries><library_visual_scenes><visual_scene id="S">'
           '<node id="a"><instance_geometry url="#G0"/></node>'
           '<node id="b"><translate>0 0 0.5</translate><instance_geometry url="#G1"/><instance_geometry url="#G2"/></node>'
           '</visual_scene></library_visual_scenes><scene><instance_visual_scene url="#S"/></scene></COLLADA>')
    p = tmp_path / "inst.dae"
    p.write_text(dae)
    hm, om, s = _both(p)
    np.testing.assert_array_equal(s["geom_ids"], [0, 0, 1])
    np.testing.assert_array_equal(s["object_ids"], [0, 1, 1])
    m = T.O1DnModel(3, 1, T.Interval(0.0, 50.0), (0.0, 0.0, 0.0),
                    np.array([[0, 0, 1], [0, 0, 1], [0, 0, 1]], np.float32))
    # rays straight up from z = 0, 1.2 and 2.7: first hit the planes at z = 1, 2.5, 3.5
    P = np.zeros((3, 8), np.float32)
    P[:, 3] = 1.0
    P[:, 6] = [0.0, 1.2, 2.7]
    g = _compare(hm, om, m, P, "instance")
    np.testing.assert_array_equal(g["geom_ids"][::3], [0, 0, 1])
    np.testing.assert_array_equal(g["object_ids"][::3], [0, 1, 1])
    np.testing.assert_allclose(g["ranges"][::3], [1.0, 1.3, 0.8], atol=1e-6)


def test_ouster_model_in_sphere():
    """The reference's OS0-128 metadata as an O1Dn model, cast inside sphere.ply: GPU vs oracle."""
    hm, om, s = _both(MAPS / "sphere.ply")
    m = o1dn_from_ouster_meta_file(MAPS.parent / "ouster" / "os0-128x1024x20-meta.json")
    P = _poses(3, -0.2, 0.2, 4)
    g = _compare(hm, om, m, P, "ouster")
    assert g["hits"].all()
