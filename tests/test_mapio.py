"""Map file loading (rmh_scene_*): formats, polygon handling, scene-graph ids and transforms.

The expected ids follow Rmagine's import (Assimp -> make_embree_scene -> freeze,
src/rmagine_embree/src/map/embree/EmbreeScene.cpp:246-290,317-406): nodes with meshes in
depth-first pre-order become instances 0..k-1, single-mesh instances are frozen into their
mesh (geom id = object id = instance id), meshes of multi-mesh instances report their index
inside the instance as geom id and the instance id as object id.  The map files under
tests/golden/maps are the reference's own data files (dat/).  Host-only: no GPU needed.
"""
import math
import struct
from pathlib import Path

import numpy as np
import pytest

from rmagine_amd import _lib
from rmagine_amd.simulation import load_map_file

MAPS = Path(__file__).resolve().parent / "golden" / "maps"


def test_reference_triangle_and_sphere():
    t = load_map_file(MAPS / "triangle.ply")
    np.testing.assert_array_equal(t["vertices"], [[0, 0, 0], [1, 0, 0], [0, 1, 0]])
    np.testing.assert_array_equal(t["faces"], [[0, 1, 2]])
    np.testing.assert_array_equal(t["mesh_face_offsets"], [0, 1])
    s = load_map_file(MAPS / "sphere.ply")
    V, F = s["vertices"], s["faces"]
    assert V.shape == (642, 3) and F.shape == (1280, 3)
    np.testing.assert_allclose(np.linalg.norm(V.astype(np.float64), axis=1), 1.0, atol=2e-6)
    # closed 2-manifold: every undirected edge in exactly two faces, consistent orientation
    edges = np.sort(np.concatenate([F[:, [0, 1]], F[:, [1, 2]], F[:, [2, 0]]]), axis=1)
    _, counts = np.unique(edges, axis=0, return_counts=True)
    assert (counts == 2).all()
    directed = np.concatenate([F[:, [0, 1]], F[:, [1, 2]], F[:, [2, 0]]])
    assert len(np.unique(directed, axis=0)) == len(directed)


def _write_ply(path, V, polys, fmt, extra_vertex_props=True):
    """PLY writer for the round-trip tests (extra properties that the loader must skip)."""
    head = ["ply", f"format {fmt} 1.0", "comment written by the test", f"element vertex {len(V)}",
            "property float x", "property double y", "property float z"]
    if extra_vertex_props:
        head += ["property uchar red", "property short s"]
    head += [f"element face {len(polys)}", "property uchar flag", "property list uchar int vertex_indices",
             "element edge 1", "property int vertex1", "property int vertex2", "end_header"]
    with open(path, "wb") as f:
        f.write(("\n".join(head) + "\n").encode())
        if fmt == "ascii":
            for v in V:
                f.write((f"{float(v[0])!r} {float(v[1])!r} {float(v[2])!r}" + (" 7 -3" if extra_vertex_props else "") + "\n").encode())
            for p in polys:
                f.write(("1 " + f"{len(p)} " + " ".join(map(str, p)) + "\n").encode())
            f.write(b"0 1\n")
        else:
            e = "<" if fmt == "binary_little_endian" else ">"
            for v in V:
                f.write(struct.pack(e + "fdf", float(v[0]), float(v[1]), float(v[2])))
                if extra_vertex_props:
                    f.write(struct.pack(e + "Bh", 7, -3))
            for p in polys:
                f.write(struct.pack(e + "BB", 1, len(p)) + struct.pack(e + f"{len(p)}i", *p))
            f.write(struct.pack(e + "ii", 0, 1))


@pytest.mark.parametrize("fmt", ["ascii", "binary_little_endian", "binary_big_endian"])
def test_ply_formats_roundtrip(tmp_path, fmt):
    rng = np.random.default_rng(5)
    V = rng.uniform(-3, 3, size=(40, 3)).astype(np.float32)
    polys = [list(rng.choice(40, 3, replace=False)) for _ in range(30)] + [[0, 1, 2, 3], [4, 5, 6, 7, 8]]
    path = tmp_path / "m.ply"
    _write_ply(path, V, polys, fmt)
    first = load_map_file(path)
    np.testing.assert_array_equal(first["vertices"], V)
    expect = np.array([p[:3] for p in polys], np.uint32)      # reference: first three indices
    np.testing.assert_array_equal(first["faces"], expect)
    fan = load_map_file(path, polygons="fan")
    expect_fan = [p[:3] for p in polys[:30]] + [[0, 1, 2], [0, 2, 3], [4, 5, 6], [4, 6, 7], [4, 7, 8]]
    np.testing.assert_array_equal(fan["faces"], np.array(expect_fan, np.uint32))


def test_obj_objects_groups_and_ids(tmp_path):
    obj = tmp_path / "scene.obj"
    obj.write_text("\n".join([
        "# two objects; the second has two material groups (a multi-mesh node)",
        "v 0 0 0", "v 1 0 0", "v 1 1 0", "v 0 1 0",
        "o first", "f 1 2 3", "f 1/1 3/1 4/1",
        "v 0 0 1", "v 1 0 1", "v 1 1 1", "v 0 1 1",
        "o second", "usemtl a", "f -4//1 -3//1 -2//1 -1//1", "f 6 7 8", "usemtl b", "f 5/1/1 7/1/1 8/1/1",
        "o third", "usemtl c", "f 1 2 3 4",
        "o empty",
    ]) + "\n")
    s = load_map_file(obj)
    # object "third" holds only a quad: no triangle, so the reference skips that mesh (its
    # node still takes instance id 2)
    np.testing.assert_array_equal(s["mesh_face_offsets"], [0, 2, 4, 5])
    np.testing.assert_array_equal(s["geom_ids"], [0, 0, 1])    # node 0 frozen; node 1: meshes 0, 1
    np.testing.assert_array_equal(s["object_ids"], [0, 1, 1])  # instance ids
    V, F = s["vertices"], s["faces"]
    np.testing.assert_array_equal(V[F[2]], [[0, 0, 1], [1, 0, 1], [1, 1, 1]])  # quad -> first triangle
    np.testing.assert_array_equal(V[F[3]], [[1, 0, 1], [1, 1, 1], [0, 1, 1]])
    np.testing.assert_array_equal(V[F[4]], [[0, 0, 1], [1, 1, 1], [0, 1, 1]])
    fan = load_map_file(obj, polygons="fan")
    np.testing.assert_array_equal(fan["mesh_face_offsets"], [0, 2, 5, 6, 8])
    np.testing.assert_array_equal(fan["geom_ids"], [0, 0, 1, 2])
    np.testing.assert_array_equal(fan["object_ids"], [0, 1, 1, 2])
    with pytest.raises(_lib.RmhError):
        bad = tmp_path / "bad.obj"
        bad.write_text("v 0 0 0\nf 1 2 3\n")
        load_map_file(bad)


def test_stl_binary_and_ascii(tmp_path):
    tris = np.array([[[0, 0, 0], [1, 0, 0], [0, 1, 0]], [[0, 0, 1], [1, 0, 1], [0, 1, 1]]], np.float32)
    b = tmp_path / "b.stl"
    with open(b, "wb") as f:
        f.write(b"\0" * 80 + struct.pack("<I", len(tris)))
        for t in tris:
            f.write(struct.pack("<3f", 0, 0, 1) + t.tobytes() + b"\0\0")
    a = tmp_path / "a.stl"
    a.write_text("solid x\n" + "".join(
        "facet normal 0 0 1\n outer loop\n" + "".join(f"  vertex {p[0]} {p[1]} {p[2]}\n" for p in t) + " endloop\nendfacet\n"
        for t in tris) + "endsolid x\n")
    for path in (b, a):
        s = load_map_file(path)
        np.testing.assert_array_equal(s["vertices"][s["faces"]], tris)


def _dae(nodes_xml, geoms_xml, unit=1.0):
    return f"""<?xml version="1.0" encoding="utf-8"?>
<COLLADA xmlns="http://www.collada.org/2005/11/COLLADASchema" version="1.4.1">
  <asset><unit name="u" meter="{unit}"/><up_axis>Z_UP</up_axis></asset>
  <library_geometries>{geoms_xml}</library_geometries>
  <library_visual_scenes><visual_scene id="S">{nodes_xml}</visual_scene></library_visual_scenes>
  <scene><instance_visual_scene url="#S"/></scene>
</COLLADA>"""


def _tri_geom(gid, pts):
    flat = " ".join(str(x) for p in pts for x in p)
    return f"""<geometry id="{gid}"><mesh>
  <source id="{gid}-pos"><float_array id="{gid}-arr" count="{3 * len(pts)}">{flat}</float_array>
    <technique_common><accessor source="#{gid}-arr" count="{len(pts)}" stride="3"/></technique_common></source>
  <vertices id="{gid}-v"><input semantic="POSITION" source="#{gid}-pos"/></vertices>
  <triangles count="1"><input semantic="VERTEX" source="#{gid}-v" offset="0"/>
    <input semantic="NORMAL" source="#{gid}-n" offset="1"/><p>0 0 1 0 2 0</p></triangles>
</mesh></geometry>"""


def test_collada_transforms_and_instances(tmp_path):
    pts = [(0, 0, 0), (1, 0, 0), (0, 1, 0)]
    geoms = _tri_geom("A", pts) + _tri_geom("B", pts)
    nodes = """
      <node id="cam" name="cam"><matrix>1 0 0 5 0 1 0 0 0 0 1 0 0 0 0 1</matrix></node>
      <node id="n0" name="n0"><translate>1 2 3</translate><rotate>0 0 1 90</rotate><scale>2 3 4</scale>
        <instance_geometry url="#A"/>
        <node id="n1" name="n1"><matrix>1 0 0 0 0 1 0 0 0 0 1 10 0 0 0 1</matrix>
          <instance_geometry url="#A"/><instance_geometry url="#B"/></node>
      </node>"""
    p = tmp_path / "s.dae"
    p.write_text(_dae(nodes, geoms, unit=0.5))
    s = load_map_file(p)
    # pre-order: n0 (single mesh -> frozen, id 0), n1 (two meshes -> instance 1)
    np.testing.assert_array_equal(s["mesh_face_offsets"], [0, 1, 2, 3])
    np.testing.assert_array_equal(s["geom_ids"], [0, 0, 1])
    np.testing.assert_array_equal(s["object_ids"], [0, 1, 1])
    U = np.diag([0.5, 0.5, 0.5, 1.0])
    T = np.eye(4); T[:3, 3] = [1, 2, 3]
    R = np.eye(4); R[:2, :2] = [[0, -1], [1, 0]]
    S = np.diag([2.0, 3.0, 4.0, 1.0])
    M0 = U @ T @ R @ S
    M1 = M0 @ np.array([[1, 0, 0, 0], [0, 1, 0, 0], [0, 0, 1, 10], [0, 0, 0, 1.0]])
    P = np.c_[np.array(pts, float), np.ones(3)]
    V, F = s["vertices"], s["faces"]
    np.testing.assert_allclose(V[F[0]], (P @ M0.T)[:, :3], atol=1e-6)
    np.testing.assert_allclose(V[F[1]], (P @ M1.T)[:, :3], atol=1e-6)
    np.testing.assert_allclose(V[F[2]], (P @ M1.T)[:, :3], atol=1e-6)


def test_collada_reference_scenes():
    box = load_map_file(MAPS / "box_rot_trans_scaled.dae")
    # node matrix of that file: 45 degrees about z, uniform scale 0.2, translation (0, -5, 0)
    c, sn = math.cos(math.pi / 4), math.sin(math.pi / 4)
    M = np.array([[0.2 * c, 0.2 * sn, 0, 0], [-0.2 * sn, 0.2 * c, 0, -5], [0, 0, 0.2, 0], [0, 0, 0, 1]])
    cube = np.array([[x, y, z, 1.0] for x in (1, -1) for y in (1, -1) for z in (1, -1)])
    got = np.unique(np.round(box["vertices"][box["faces"]].reshape(-1, 3).astype(np.float64), 5), axis=0)
    want = np.unique(np.round((cube @ M.T)[:, :3], 5), axis=0)
    np.testing.assert_allclose(got, want, atol=2e-5)
    two = load_map_file(MAPS / "two_cubes.dae")
    np.testing.assert_array_equal(two["geom_ids"], [0, 1])
    np.testing.assert_array_equal(two["object_ids"], [0, 1])
    many = load_map_file(MAPS / "many_objects.dae")
    n = len(many["geom_ids"])
    np.testing.assert_array_equal(many["geom_ids"], np.arange(n))
    np.testing.assert_array_equal(many["object_ids"], np.arange(n))
    assert len(many["faces"]) == many["mesh_face_offsets"][-1] == 2331


def test_errors(tmp_path):
    with pytest.raises(_lib.RmhError):
        load_map_file(tmp_path / "missing.ply")
    g = tmp_path / "x.xyz"
    g.write_text("nothing")
    with pytest.raises(_lib.RmhError):
        load_map_file(g)
    with pytest.raises(ValueError):
        load_map_file(MAPS / "triangle.ply", polygons="strip")
    q = tmp_path / "quads.obj"
    q.write_text("v 0 0 0\nv 1 0 0\nv 1 1 0\nv 0 1 0\nf 1 2 3 4\n")
    with pytest.raises(_lib.RmhError):  # a mesh without triangles is skipped -> nothing left
        load_map_file(q)
    assert len(load_map_file(q, polygons="fan")["faces"]) == 2
