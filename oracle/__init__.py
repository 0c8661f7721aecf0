"""ORACLE -- test infrastructure, not product code.

Python entry points of the CPU oracle.  Only tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg import this package, and only as the checker (or the timed
CPU baseline).  The product (rmagine_amd / librmh.so) never imports it.

  simulate      -> oracle/oracle.cpp (restates SphereSimulatorEmbree.tcc:50-254 and
                   its Pinhole/O1Dn/OnDn siblings; see the header of that file)
  statistics    -> oracle/stats.py (restates statistics.cpp:20-126, CrossStatistics.tcc:9-58,
                   optimization.cpp:12-55 in FP64)

Parity pins are listed in each file's header.  The shared library is built by
`make -C oracle` (called from __graft_entry__.build()).
"""
from __future__ import annotations

import ctypes
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
LIB_PATH = HERE / "_build" / "liboracle.so"

_lib = None

FLAG_EDGE = 1       # hit within 1e-6 (barycentric) of a triangle edge
FLAG_NEAR_MISS = 2  # another triangle missed by < 1e-6 at a competitive distance
FLAG_TIE = 4        # a second hit within a relative 1e-6 in distance
FLAG_RANGE = 8      # distance within a relative 1e-6 of range.min / range.max


def build() -> Path:
    import subprocess
    subprocess.run(["make", "-s", "-C", str(HERE), "-B"], check=True)
    return LIB_PATH


def lib():
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            build()
        _lib = ctypes.CDLL(str(LIB_PATH))
        P = ctypes.c_void_p
        _lib.oracle_map_build.restype = P
        _lib.oracle_map_build.argtypes = [P, ctypes.c_uint32, P, ctypes.c_uint32, P, ctypes.c_uint32]
        _lib.oracle_map_free.restype = None
        _lib.oracle_map_free.argtypes = [P]
        _lib.oracle_map_set_mesh_ids.restype = None
        _lib.oracle_map_set_mesh_ids.argtypes = [P, P, P]
        _lib.oracle_simulate.restype = ctypes.c_int
        _lib.oracle_simulate.argtypes = [P, ctypes.c_int, P, P, P, P, P, ctypes.c_uint32,
                                         P, P, P, P, P, P, P, P, ctypes.c_int, ctypes.c_int]
    return _lib


def _p(a):
    return ctypes.c_void_p(0) if a is None else a.ctypes.data_as(ctypes.c_void_p)


class OracleMap:
    def __init__(self, V, F, mesh_face_offsets=None, geom_ids=None, object_ids=None):
        self.V = np.ascontiguousarray(np.asarray(V, dtype=np.float32).reshape(-1, 3))
        self.F = np.ascontiguousarray(np.asarray(F, dtype=np.uint32).reshape(-1, 3))
        off = None
        n_meshes = 0
        if mesh_face_offsets is not None:
            off = np.ascontiguousarray(np.asarray(mesh_face_offsets, dtype=np.uint32))
            n_meshes = len(off) - 1
        self._h = lib().oracle_map_build(_p(self.V), len(self.V), _p(self.F), len(self.F), _p(off), n_meshes)
        if geom_ids is not None or object_ids is not None:
            g = None if geom_ids is None else np.ascontiguousarray(np.asarray(geom_ids, dtype=np.uint32))
            o = None if object_ids is None else np.ascontiguousarray(np.asarray(object_ids, dtype=np.uint32))
            lib().oracle_map_set_mesh_ids(self._h, _p(g), _p(o))

    def __del__(self):
        if getattr(self, "_h", None):
            lib().oracle_map_free(self._h)
            self._h = None


ATTRS = ("hits", "ranges", "points", "normals", "face_ids", "geom_ids", "object_ids")


def simulate(omap: OracleMap, model, Tbm, Tsb=None, attrs=ATTRS, threads: int = 0, brute: bool = False,
             flags: bool = True) -> dict:
    """Embree-semantics simulation of `model` from every pose in Tbm ((N,8) float32)."""
    from rmagine_amd import types as T  # model definitions only

    Tbm = np.ascontiguousarray(np.asarray(Tbm, dtype=np.float32).reshape(-1, 8))
    Tsb = T.identity_transforms(1)[0] if Tsb is None else np.asarray(Tsb, dtype=np.float32).reshape(8)
    Tsb = np.ascontiguousarray(Tsb)
    n = len(Tbm) * model.getWidth() * model.getHeight()
    out = {}
    shapes = {"hits": (np.uint8, (n,)), "ranges": (np.float32, (n,)), "points": (np.float32, (n, 3)),
              "normals": (np.float32, (n, 3)), "face_ids": (np.uint32, (n,)), "geom_ids": (np.uint32, (n,)),
              "object_ids": (np.uint32, (n,))}
    for a in attrs:
        dt, sh = shapes[a]
        out[a] = np.empty(sh, dtype=dt)
    fl = np.empty(n, dtype=np.uint8) if flags else None
    mf = T.model_to_floats(model)
    dirs = origs = None
    if isinstance(model, (T.O1DnModel, T.OnDnModel)):
        dirs = np.ascontiguousarray(np.asarray(model.dirs, dtype=np.float32).reshape(-1, 3))
    if isinstance(model, T.OnDnModel):
        origs = np.ascontiguousarray(np.asarray(model.origs, dtype=np.float32).reshape(-1, 3))
    lib().oracle_simulate(omap._h, model.kind, _p(mf), _p(dirs), _p(origs), _p(Tsb), _p(Tbm), len(Tbm),
                          _p(out.get("hits")), _p(out.get("ranges")), _p(out.get("points")),
                          _p(out.get("normals")), _p(out.get("face_ids")), _p(out.get("geom_ids")),
                          _p(out.get("object_ids")), _p(fl), int(threads), int(bool(brute)))
    if flags:
        out["flags"] = fl
    return out
