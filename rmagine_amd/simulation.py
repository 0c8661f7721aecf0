"""Python mirror of Rmagine's GPU Simulator API on top of librmh.

Mirrors (uos/rmagine v2.4.2):
  SimulatorOptix: setMap / setTsb              src/rmagine_optix/include/rmagine/simulation/SimulatorOptix.hpp:22-79
  SphereSimulatorOptix: setModel / simulate     src/rmagine_optix/include/rmagine/simulation/SphereSimulatorOptix.hpp:122-191
  O1DnSimulatorOptix::setModel (uploads dirs)   src/rmagine_optix/src/simulation/O1DnSimulatorOptix.cpp:41-83
  OptixMap / import_optix_map                   src/rmagine_optix/include/rmagine/map/OptixMap.hpp:88-110
  Bundle + resize_memory_bundle                 src/rmagine_core/include/rmagine/simulation/SimulationResults.hpp:176-216

Device buffers are torch tensors on a HIP device (torch is used for device memory and
streams only; every computation runs in librmh).  Transforms are float32 tensors of
shape (N, 8).  Id attributes are int32 tensors holding the uint32 bit patterns
(0xFFFFFFFF reads as -1).
"""
from __future__ import annotations

import ctypes

import numpy as np
import torch

from . import _lib
from . import types as T

IntAttrAll = T.ATTRIBUTES

_TORCH_DTYPES = {
    "hits": torch.uint8,
    "ranges": torch.float32,
    "points": torch.float32,
    "normals": torch.float32,
    "face_ids": torch.int32,
    "geom_ids": torch.int32,
    "object_ids": torch.int32,
}


def _ptr(t) -> ctypes.c_void_p:
    return ctypes.c_void_p(0 if t is None else int(t.data_ptr()))


def stream_handle(stream=None) -> ctypes.c_void_p:
    s = stream if stream is not None else torch.cuda.current_stream()
    return ctypes.c_void_p(int(s.cuda_stream))


def _device_tensor(x, dtype, device) -> torch.Tensor:
    if isinstance(x, torch.Tensor):
        return x.to(device=device, dtype=dtype).contiguous()
    return torch.as_tensor(np.ascontiguousarray(x), dtype=dtype, device=device)


class Bundle:
    """Result buffers; an attribute is produced iff its tensor has size > 0."""

    def __init__(self, **attrs):
        for name in T.ATTRIBUTES:
            setattr(self, name, attrs.get(name))

    def enabled(self):
        return [n for n in T.ATTRIBUTES if getattr(self, n) is not None and getattr(self, n).numel() > 0]

    def to_numpy(self) -> dict:
        out = {}
        for n in self.enabled():
            a = getattr(self, n).cpu().numpy()
            if T.ATTR_DTYPES[n][0] == np.uint32:
                a = a.view(np.uint32)
            out[n] = a
        return out


def resize_memory_bundle(bundle: Bundle, W: int, H: int, N: int = 1, attrs=IntAttrAll, device="cuda") -> Bundle:
    """Allocates W*H*N entries for each requested attribute (SimulationResults.hpp:176-216)."""
    n = int(W) * int(H) * int(N)
    for name in attrs:
        comp = T.ATTR_DTYPES[name][1]
        shape = (n,) if comp == 1 else (n, comp)
        setattr(bundle, name, torch.empty(shape, dtype=_TORCH_DTYPES[name], device=device))
    return bundle


class HipMap:
    """Mesh plus BVH resident in HBM; the role of OptixMap/EmbreeMap.

    vertices: (V, 3) float32, faces: (F, 3) uint32 indices.  mesh_face_offsets (len
    n_meshes+1) splits the faces into separate geometries (geom ids 0..n_meshes-1).
    wide=True also collapses the binary BVH into the 4-wide tree with 8-bit quantised
    child boxes and makes the ray kernels traverse it (same results; the binary default is
    faster on the measured workloads); set_layout() switches later.  refine=False skips the
    SAH refinement (treelet restructuring) of the radix tree: faster build, slower rays.
    """

    def __init__(self, vertices, faces, mesh_face_offsets=None, max_leaf_size: int = 4,
                 sah_cost_node: float = 1.0, sah_cost_tri: float = 1.5, device="cuda", stream=None,
                 wide: bool = False, refine: bool = True):
        V = _device_tensor(np.asarray(vertices, dtype=np.float32).reshape(-1, 3), torch.float32, device)
        F = _device_tensor(np.asarray(faces, dtype=np.uint32).reshape(-1, 3).view(np.int32), torch.int32, device)
        self.device = torch.device(device)
        self._handle = ctypes.c_void_p()
        off_arr, n_meshes = None, 0
        if mesh_face_offsets is not None:
            off = np.ascontiguousarray(np.asarray(mesh_face_offsets, dtype=np.uint32))
            n_meshes = len(off) - 1
            off_arr = (ctypes.c_uint32 * len(off))(*off.tolist())
        prm = T.rmh_bvh_params(max_leaf_size, sah_cost_node, sah_cost_tri,
                               (T.BVH_WIDE if wide else 0) | (0 if refine else T.BVH_NO_REFINE))
        _lib.call("rmh_map_build", _ptr(V), V.shape[0], _ptr(F), F.shape[0], off_arr, n_meshes,
                  ctypes.byref(prm), stream_handle(stream), ctypes.byref(self._handle))
        self.n_vertices = V.shape[0]
        self.n_faces = F.shape[0]

    @property
    def handle(self) -> ctypes.c_void_p:
        return self._handle

    def refit(self, vertices, stream=None) -> None:
        """New positions for the map's vertices ((V,3), same faces): triangles, normals and
        BVH boxes are updated in place without rebuilding the tree."""
        Vt = _device_tensor(np.asarray(vertices, dtype=np.float32).reshape(-1, 3) if not isinstance(vertices, torch.Tensor)
                            else vertices, torch.float32, self.device)
        if Vt.shape != (self.n_vertices, 3):
            raise ValueError(f"refit needs ({self.n_vertices}, 3) vertices")
        _lib.call("rmh_map_refit", self._handle, _ptr(Vt), stream_handle(stream))

    def set_layout(self, wide: bool, stream=None) -> bool:
        """Node layout the ray kernels traverse: the 4-wide quantised tree (wide=True) or the
        binary one.  Returns whether the 4-wide tree is in use (a tree it cannot express stays
        binary).  No launch on this map may be in flight; re-capture graphs afterwards."""
        _lib.call("rmh_map_set_layout", self._handle, 1 if wide else 0, stream_handle(stream))
        return self.info()["n_wide_nodes"] > 0

    @property
    def wide(self) -> bool:
        return self.info()["n_wide_nodes"] > 0

    def set_mesh_ids(self, geom_ids=None, object_ids=None) -> None:
        """Geometry / object id reported for each mesh (defaults: the mesh index)."""
        n = self.info()["n_meshes"]

        def arr(x):
            if x is None:
                return None
            a = np.ascontiguousarray(np.asarray(x, dtype=np.uint32))
            if a.shape != (n,):
                raise ValueError(f"expected {n} ids")
            return (ctypes.c_uint32 * n)(*a.tolist())
        _lib.call("rmh_map_set_mesh_ids", self._handle, arr(geom_ids), arr(object_ids), n)

    def closest_points(self, queries: torch.Tensor, max_distance: float = float(np.finfo(np.float32).max),
                       stream=None) -> dict:
        """Nearest map point of every query ((N,3) float32 device tensor) strictly within
        max_distance: the batched form of EmbreeMap::closestPoint
        (src/rmagine_embree/src/map/EmbreeMap.cpp:33-37).  Returns dist (N,), points (N,3),
        normals (N,3), face_ids, geom_ids, object_ids (N,) device tensors; misses have dist =
        FLT_MAX, ids 0xFFFFFFFF (as int32 -1) and NaN point/normal.  Batches of 65,536
        queries or more are sorted on the device by the Morton code of their positions and
        walked in that order when their own order is spatially incoherent; the results are
        per query and do not depend on the order."""
        if not (isinstance(queries, torch.Tensor) and queries.dtype == torch.float32 and queries.is_cuda
                and queries.is_contiguous() and queries.dim() == 2 and queries.shape[1] == 3):
            raise ValueError("queries must be a contiguous (N,3) float32 CUDA tensor")
        n = queries.shape[0]
        dev = queries.device
        res = {"dist": torch.empty(n, dtype=torch.float32, device=dev),
               "points": torch.empty((n, 3), dtype=torch.float32, device=dev),
               "normals": torch.empty((n, 3), dtype=torch.float32, device=dev),
               "face_ids": torch.empty(n, dtype=torch.int32, device=dev),
               "geom_ids": torch.empty(n, dtype=torch.int32, device=dev),
               "object_ids": torch.empty(n, dtype=torch.int32, device=dev)}
        out = T.rmh_closest_outputs(*(res[k].data_ptr() for k in ("dist", "points", "normals", "face_ids",
                                                                   "geom_ids", "object_ids")))
        _lib.call("rmh_closest_points", self._handle, ctypes.c_void_p(queries.data_ptr()), n,
                  ctypes.c_float(max_distance), ctypes.byref(out), stream_handle(stream))
        return res

    def info(self) -> dict:
        inf = T.rmh_map_info()
        _lib.call("rmh_map_get_info", self._handle, ctypes.byref(inf))
        return {k: getattr(inf, k) for k, _ in inf._fields_}

    def __del__(self):
        h = getattr(self, "_handle", None)
        if h is not None and h.value:
            try:
                _lib.lib().rmh_map_destroy(h)
            except Exception:
                pass
            self._handle = ctypes.c_void_p()


def load_map_file(path, polygons: str = "first") -> dict:
    """Reads a mesh file (PLY, OBJ, STL, COLLADA) into flat host arrays, with the scene graph
    resolved as Rmagine's import does (import_embree_map: Assimp, make_embree_scene, freeze;
    src/rmagine_embree/include/rmagine/map/EmbreeMap.hpp:94-117).

    polygons: "first" keeps the first three indices of every polygon (the reference's
    EmbreeMesh::init), "fan" triangulates polygons as fans.
    Returns vertices (V,3) float32, faces (F,3) uint32, mesh_face_offsets (M+1,), geom_ids
    and object_ids (M,) uint32.
    """
    if polygons not in ("first", "fan"):
        raise ValueError("polygons must be 'first' or 'fan'")
    flags = 1 if polygons == "fan" else 0
    h = ctypes.c_void_p()
    _lib.call("rmh_scene_load", str(path).encode(), flags, ctypes.byref(h))
    try:
        nv, nf, nm = ctypes.c_uint32(), ctypes.c_uint32(), ctypes.c_uint32()
        _lib.call("rmh_scene_counts", h, ctypes.byref(nv), ctypes.byref(nf), ctypes.byref(nm))
        out = {"vertices": np.empty((nv.value, 3), np.float32), "faces": np.empty((nf.value, 3), np.uint32),
               "mesh_face_offsets": np.empty(nm.value + 1, np.uint32), "geom_ids": np.empty(nm.value, np.uint32),
               "object_ids": np.empty(nm.value, np.uint32)}
        _lib.call("rmh_scene_copy", h, *(out[k].ctypes.data for k in
                                         ("vertices", "faces", "mesh_face_offsets", "geom_ids", "object_ids")))
    finally:
        _lib.lib().rmh_scene_destroy(h)
    return out


def import_hip_map(path, polygons: str = "first", device="cuda", **bvh) -> HipMap:
    """HipMap of a mesh file: the counterpart of import_embree_map / import_optix_map
    (src/rmagine_optix/include/rmagine/map/OptixMap.hpp:88-109).  Hits report the geometry
    and object ids Rmagine's Embree scene would give the same file."""
    s = load_map_file(path, polygons)
    m = HipMap(s["vertices"], s["faces"], s["mesh_face_offsets"], device=device, **bvh)
    m.set_mesh_ids(s["geom_ids"], s["object_ids"])
    return m


RAY_ORDERS = ("list", "sorted", "tiles")  # rmh.h RMH_RAY_ORDER_LIST / _SORTED / _TILES


def _ray_order_code(order) -> int:
    """True / "sorted": by direction; False / "list": list order; "tiles": the W x H grid in
    4 x 16 tiles."""
    if isinstance(order, str):
        if order not in RAY_ORDERS:
            raise ValueError(f"ray order must be one of {RAY_ORDERS}")
        return RAY_ORDERS.index(order)
    return 1 if order else 0


class _SensorHandle:
    def __init__(self, model, device):
        self.model = model
        keep = []
        dirs_dev = origs_dev = 0
        if isinstance(model, (T.O1DnModel, T.OnDnModel)):
            d = _device_tensor(np.asarray(model.dirs, dtype=np.float32).reshape(-1, 3), torch.float32, device)
            keep.append(d)
            dirs_dev = int(d.data_ptr())
            if isinstance(model, T.OnDnModel):
                o = _device_tensor(np.asarray(model.origs, dtype=np.float32).reshape(-1, 3), torch.float32, device)
                keep.append(o)
                origs_dev = int(o.data_ptr())
            if d.shape[0] != model.width * model.height:
                raise ValueError("dirs must hold width*height directions")
        kind, cmodel = T.model_to_c(model, dirs_dev, origs_dev)
        self.handle = ctypes.c_void_p()
        _lib.call("rmh_sensor_create", kind, ctypes.byref(cmodel), ctypes.byref(self.handle))
        self.W, self.H = model.getWidth(), model.getHeight()

    def set_ray_order(self, order) -> None:
        _lib.call("rmh_sensor_set_ray_order", self.handle, _ray_order_code(order))

    def ray_order(self) -> str:
        v = ctypes.c_int()
        _lib.call("rmh_sensor_get_ray_order", self.handle, ctypes.byref(v))
        return RAY_ORDERS[v.value]

    def ray_order_sorted(self) -> bool:
        return self.ray_order() == "sorted"

    def set_walk(self, speculative: bool) -> None:
        _lib.call("rmh_sensor_set_walk", self.handle, 1 if speculative else 0)

    def walk_speculative(self) -> bool:
        v = ctypes.c_int()
        _lib.call("rmh_sensor_get_walk", self.handle, ctypes.byref(v))
        return bool(v.value)

    def __del__(self):
        h = getattr(self, "handle", None)
        if h is not None and h.value:
            try:
                _lib.lib().rmh_sensor_destroy(h)
            except Exception:
                pass
            self.handle = ctypes.c_void_p()


class SimulatorHip:
    """Common part of the *SimulatorHip classes (SimulatorOptix.hpp:22-79)."""

    _model_type = None

    def __init__(self, map: HipMap | None = None, device="cuda"):
        self.device = torch.device(device)
        self._map = map
        self._sensor = None
        self._Tsb = T.identity_transforms(1)[0]

    def setMap(self, map: HipMap) -> None:
        self._map = map

    def map(self):
        return self._map

    def setTsb(self, Tsb) -> None:
        self._Tsb = np.asarray(Tsb, dtype=np.float32).reshape(8).copy()

    def setModel(self, model) -> None:
        if self._model_type is not None and not isinstance(model, self._model_type):
            raise TypeError(f"{type(self).__name__} expects {self._model_type.__name__}")
        self._sensor = _SensorHandle(model, self.device)

    @property
    def model(self):
        return None if self._sensor is None else self._sensor.model

    def set_ray_order(self, sorted_=True) -> None:
        """O1Dn/OnDn: cast the rays in direction-sorted order (True or "sorted"), in list
        order (False or "list"), or as the model's W x H grid in 4 x 16 tiles ("tiles", as
        image sensors are cast).  Default: "tiles" for models of at least 4 rows and 16
        columns, else "sorted".  The results are identical; only the speed differs.  Image
        models ignore it."""
        if self._sensor is None:
            raise RuntimeError("set_ray_order: no sensor model set")
        self._sensor.set_ray_order(sorted_)

    def ray_order(self) -> str:
        return "list" if self._sensor is None else self._sensor.ray_order()

    def ray_order_sorted(self) -> bool:
        return self._sensor is not None and self._sensor.ray_order_sorted()

    def set_walk(self, speculative: bool = True) -> None:
        """O1Dn/OnDn: walk the BVH speculatively (lanes that reach a leaf hold it while the
        wave walks on; fastest for scattered rays, the default of lists of fewer than 4 rows
        or 16 columns) or the classic way (the wave walks until every lane holds a leaf;
        fastest for a scanner's coherent ray grid, the default of larger grids).  The results
        are identical; only the speed differs.  Image models ignore it."""
        if self._sensor is None:
            raise RuntimeError("set_walk: no sensor model set")
        self._sensor.set_walk(speculative)

    def walk_speculative(self) -> bool:
        return self._sensor is not None and self._sensor.walk_speculative()

    def simulate(self, Tbm, bundle: Bundle | None = None, attrs=IntAttrAll, stream=None) -> Bundle:
        """Casts every ray of the model from every pose in Tbm ((N,8) float32 tensor).

        If bundle is None a new one is allocated with `attrs` (resize_memory_bundle);
        otherwise its non-empty tensors receive the results.  Asynchronous on `stream`.
        """
        if self._map is None:
            raise _lib.NoMapError(3, "simulate: no map set")  # SphereSimulatorOptix.tcc:49-53
        if self._sensor is None:
            raise RuntimeError("simulate: no sensor model set")
        if not isinstance(Tbm, torch.Tensor):
            Tbm = torch.as_tensor(np.asarray(Tbm, dtype=np.float32).reshape(-1, 8), device=self.device)
        if Tbm.dtype != torch.float32 or Tbm.dim() != 2 or Tbm.shape[1] != 8 or not Tbm.is_contiguous():
            raise ValueError("Tbm must be a contiguous float32 tensor of shape (N, 8)")
        if Tbm.device.type != "cuda":
            raise ValueError("Tbm must be a device tensor")
        N = Tbm.shape[0]
        W, H = self._sensor.W, self._sensor.H
        if bundle is None:
            bundle = resize_memory_bundle(Bundle(), W, H, N, attrs, device=self.device)
        need = W * H * N
        outs = T.rmh_outputs()
        for name in T.ATTRIBUTES:
            t = getattr(bundle, name)
            if t is None or t.numel() == 0:
                continue
            comp = T.ATTR_DTYPES[name][1]
            if t.numel() != need * comp or not t.is_contiguous() or t.device.type != "cuda":
                raise ValueError(f"bundle.{name} must be a contiguous device tensor of {need}x{comp} elements")
            if t.dtype != _TORCH_DTYPES[name]:
                raise ValueError(f"bundle.{name} must be {_TORCH_DTYPES[name]}")
            setattr(outs, name, int(t.data_ptr()))
        Tsb = T.transform_to_c(self._Tsb)
        _lib.call("rmh_simulate", self._map.handle, self._sensor.handle, ctypes.byref(Tsb), _ptr(Tbm), N,
                  ctypes.byref(outs), stream_handle(stream))
        return bundle


    def visit_stats(self, Tbm, stream=None) -> dict:
        """Diagnostic: traversal counts of simulate()'s kernel over the same rays (binary
        nodes), summed over all rays: node visits through vector loads and through the scalar
        cache, leaves, triangle tests, wave iterations of the node loop and how many of them
        were wave-uniform; plus the per-ray means."""
        if self._map is None:
            raise _lib.NoMapError(3, "visit_stats: no map set")
        if self._sensor is None:
            raise RuntimeError("visit_stats: no sensor model set")
        Tbm = torch.as_tensor(np.asarray(Tbm, dtype=np.float32).reshape(-1, 8), device=self.device) \
            if not isinstance(Tbm, torch.Tensor) else Tbm
        if Tbm.dtype != torch.float32 or Tbm.dim() != 2 or Tbm.shape[1] != 8 or not Tbm.is_contiguous():
            raise ValueError("Tbm must be a contiguous float32 tensor of shape (N, 8)")
        counts = (ctypes.c_uint64 * 6)()
        Tsb = T.transform_to_c(self._Tsb)
        _lib.call("rmh_simulate_visit_stats", self._map.handle, self._sensor.handle, ctypes.byref(Tsb), _ptr(Tbm),
                  Tbm.shape[0], counts, stream_handle(stream))
        names = ("vector_node_visits", "scalar_node_visits", "leaves", "triangle_tests", "wave_iterations",
                 "uniform_wave_iterations")
        out = {n: int(c) for n, c in zip(names, counts)}
        rays = Tbm.shape[0] * self._sensor.W * self._sensor.H
        out["rays"] = rays
        for n in names[:4]:
            out[n + "_per_ray"] = out[n] / rays if rays else 0.0
        return out


class SphereSimulatorHip(SimulatorHip):
    _model_type = T.SphericalModel


class PinholeSimulatorHip(SimulatorHip):
    _model_type = T.PinholeModel


class O1DnSimulatorHip(SimulatorHip):
    _model_type = T.O1DnModel


class OnDnSimulatorHip(SimulatorHip):
    _model_type = T.OnDnModel


def simulator_for(model, map: HipMap | None = None, device="cuda") -> SimulatorHip:
    """make_simulator analogue (src/rmagine_core/include/rmagine/simulation/Simulator.hpp:52-78)."""
    cls = {T.SphericalModel: SphereSimulatorHip, T.PinholeModel: PinholeSimulatorHip,
           T.O1DnModel: O1DnSimulatorHip, T.OnDnModel: OnDnSimulatorHip}[type(model)]
    sim = cls(map, device=device)
    sim.setModel(model)
    return sim


class BatchSimulatorHip:
    """Many sensors -- any mix of models, poses and ray counts -- cast in ONE persistent
    launch (launches=2 splits it into image sensors, then ray lists).

    entries: list of dicts {"model": sensor model, "Tbm": (N,8) device tensor,
    "attrs": attribute names (default hits+ranges), "Tsb": optional (8,) pose}.
    After simulate(), self.bundles[i] holds entry i's results (the same layout rmh_simulate
    gives for that sensor alone).  The device table is built once; launches can be captured
    into a HipGraph.  The reference has no batched form (one optixLaunch per model).
    """

    def __init__(self, map: HipMap, entries, chunk_rays: int = 0, waves_per_cu: int = 0, tile_rows: int = 0,
                 device="cuda", lds_stack: int = 0, launches: int = 0):
        self._map = map
        self.device = torch.device(device)
        self._sensors, self._keep, self.bundles = [], [], []
        table = (T.rmh_batch_entry * max(1, len(entries)))()
        self.total_rays = 0
        for i, e in enumerate(entries):
            sh = _SensorHandle(e["model"], self.device)
            Tbm = e["Tbm"]
            if not (isinstance(Tbm, torch.Tensor) and Tbm.dtype == torch.float32 and Tbm.is_contiguous()
                    and Tbm.dim() == 2 and Tbm.shape[1] == 8 and Tbm.device.type == "cuda"):
                raise ValueError("entry Tbm must be a contiguous float32 (N,8) device tensor")
            n = Tbm.shape[0]
            b = resize_memory_bundle(Bundle(), sh.W, sh.H, n, e.get("attrs", ("hits", "ranges")), device=self.device)
            outs = T.rmh_outputs()
            for name in b.enabled():
                setattr(outs, name, int(getattr(b, name).data_ptr()))
            Tsb = e.get("Tsb")
            Tsb = T.identity_transforms(1)[0] if Tsb is None else np.asarray(Tsb, np.float32)
            table[i] = T.rmh_batch_entry(sh.handle.value, T.transform_to_c(Tsb), int(Tbm.data_ptr()), n, 0, outs)
            self._sensors.append(sh)
            self._keep.append(Tbm)
            self.bundles.append(b)
            self.total_rays += sh.W * sh.H * n
        prm = T.rmh_batch_params(chunk_rays, waves_per_cu, tile_rows, lds_stack, launches)
        self._h = ctypes.c_void_p()
        _lib.call("rmh_batch_create", map.handle, table, len(entries), ctypes.byref(prm), ctypes.byref(self._h))

    @property
    def map(self) -> HipMap:
        return self._map

    def set_params(self, chunk_rays: int = 0, waves_per_cu: int = 0, tile_rows: int = 0,
                   lds_stack: int = 0, launches: int = 0) -> None:
        """Launch shape: rays per work grab, persistent waves per CU, image rows per 64-ray
        tile, traversal-stack entries held in LDS (8, 10 or 16; deeper entries spill to
        private memory), persistent launches (1: every entry; 2: image sensors, then ray
        lists) -- 0 keeps the current value."""
        prm = T.rmh_batch_params(chunk_rays, waves_per_cu, tile_rows, lds_stack, launches)
        _lib.call("rmh_batch_set_params", self._h, ctypes.byref(prm))

    def set_ray_order(self, sorted_: bool) -> None:
        """Ray order of every O1Dn/OnDn entry (direction-sorted or list order); results are
        identical either way.  Not while a launch of this batch is in flight."""
        for sh in self._sensors:
            sh.set_ray_order(sorted_)

    def ray_order_sorted(self) -> bool:
        return any(sh.ray_order_sorted() for sh in self._sensors)

    def params(self) -> dict:
        prm = T.rmh_batch_params()
        rays = ctypes.c_uint64()
        chunks = ctypes.c_uint32()
        _lib.call("rmh_batch_get_info", self._h, ctypes.byref(rays), ctypes.byref(chunks), ctypes.byref(prm))
        return {"chunk_rays": prm.chunk_rays, "waves_per_cu": prm.waves_per_cu, "tile_rows": prm.tile_rows,
                "lds_stack": prm.lds_stack, "launches": prm.launches, "total_rays": rays.value,
                "total_chunks": chunks.value}

    def simulate(self, stream=None) -> list:
        _lib.call("rmh_batch_simulate", self._h, stream_handle(stream))
        return self.bundles

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and h.value:
            try:
                _lib.lib().rmh_batch_destroy(h)
            except Exception:
                pass
            self._h = ctypes.c_void_p()


class HipGraph:
    """A recorded sequence of librmh launches on one stream, replayed with one call.

    Usage: g = HipGraph.capture(lambda s: ..., stream) with a non-default torch stream; the
    captured function must not allocate (pre-allocate outputs and workspaces).
    """

    def __init__(self, handle):
        self._h = handle

    @classmethod
    def capture(cls, fn, stream: torch.cuda.Stream):
        if stream is None or int(stream.cuda_stream) == 0:
            raise ValueError("capture needs a non-default stream")
        torch.cuda.synchronize()
        _lib.call("rmh_graph_begin_capture", stream_handle(stream))
        h = ctypes.c_void_p()
        try:
            fn(stream)
        finally:
            _lib.call("rmh_graph_end_capture", stream_handle(stream), ctypes.byref(h))
        return cls(h)

    def launch(self, stream: torch.cuda.Stream) -> None:
        _lib.call("rmh_graph_launch", self._h, stream_handle(stream))

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and h.value:
            try:
                _lib.lib().rmh_graph_destroy(h)
            except Exception:
                pass
            self._h = ctypes.c_void_p()
