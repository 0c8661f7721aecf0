"""ctypes binding of librmh.so (include/rmh.h).

The library is loaded from rmagine_amd/_lib/librmh.so.  There is no fallback: if the
library is missing every entry point raises, so nothing silently runs on the CPU.
"""
from __future__ import annotations

import ctypes
import os
from pathlib import Path

from . import types as T

LIB_PATH = Path(__file__).resolve().parent / "_lib" / "librmh.so"

_lib = None


class RmhError(RuntimeError):
    """Error reported by librmh (status code + rmh_last_error message)."""

    def __init__(self, code: int, msg: str):
        super().__init__(f"librmh error {code}: {msg}")
        self.code = code


class NoMapError(RmhError):
    """simulate() without a map (the reference throws std::runtime_error)."""


c_void_p = ctypes.c_void_p
c_u32 = ctypes.c_uint32
c_int = ctypes.c_int
c_size = ctypes.c_size_t

_SIGNATURES = {
    "rmh_last_error": (ctypes.c_char_p, []),
    "rmh_version": (c_int, []),
    "rmh_get_device_count": (c_int, [ctypes.POINTER(c_int)]),
    "rmh_set_device": (c_int, [c_int]),
    "rmh_malloc": (c_int, [ctypes.POINTER(c_void_p), c_size]),
    "rmh_free": (c_int, [c_void_p]),
    "rmh_memcpy": (c_int, [c_void_p, c_void_p, c_size, c_int]),
    "rmh_memset": (c_int, [c_void_p, c_int, c_size]),
    "rmh_stream_create": (c_int, [ctypes.POINTER(c_void_p)]),
    "rmh_stream_destroy": (c_int, [c_void_p]),
    "rmh_stream_synchronize": (c_int, [c_void_p]),
    "rmh_device_synchronize": (c_int, []),
    "rmh_map_build": (c_int, [c_void_p, c_u32, c_void_p, c_u32, ctypes.POINTER(c_u32), c_u32,
                              ctypes.POINTER(T.rmh_bvh_params), c_void_p, ctypes.POINTER(c_void_p)]),
    "rmh_map_destroy": (c_int, [c_void_p]),
    "rmh_map_get_info": (c_int, [c_void_p, ctypes.POINTER(T.rmh_map_info)]),
    "rmh_map_refit": (c_int, [c_void_p, c_void_p, c_void_p]),
    "rmh_map_set_layout": (c_int, [c_void_p, c_int, c_void_p]),
    "rmh_map_set_mesh_ids": (c_int, [c_void_p, ctypes.POINTER(c_u32), ctypes.POINTER(c_u32), c_u32]),
    "rmh_scene_load": (c_int, [ctypes.c_char_p, c_u32, ctypes.POINTER(c_void_p)]),
    "rmh_scene_counts": (c_int, [c_void_p, ctypes.POINTER(c_u32), ctypes.POINTER(c_u32), ctypes.POINTER(c_u32)]),
    "rmh_scene_copy": (c_int, [c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p]),
    "rmh_scene_destroy": (c_int, [c_void_p]),
    "rmh_closest_points": (c_int, [c_void_p, c_void_p, c_u32, ctypes.c_float, ctypes.POINTER(T.rmh_closest_outputs),
                                   c_void_p]),
    "rmh_ouster_o1dn": (c_int, [ctypes.c_char_p, ctypes.POINTER(c_u32), ctypes.POINTER(c_u32), c_void_p, c_size]),
    "rmh_noise_apply": (c_int, [ctypes.POINTER(T.rmh_noise_params), c_void_p, ctypes.c_uint64, c_u32, c_void_p]),
    "rmh_sensor_create": (c_int, [c_int, c_void_p, ctypes.POINTER(c_void_p)]),
    "rmh_sensor_destroy": (c_int, [c_void_p]),
    "rmh_sensor_dims": (c_int, [c_void_p, ctypes.POINTER(c_u32), ctypes.POINTER(c_u32)]),
    "rmh_sensor_set_ray_order": (c_int, [c_void_p, c_int]),
    "rmh_sensor_get_ray_order": (c_int, [c_void_p, ctypes.POINTER(c_int)]),
    "rmh_sensor_set_walk": (c_int, [c_void_p, c_int]),
    "rmh_sensor_get_walk": (c_int, [c_void_p, ctypes.POINTER(c_int)]),
    "rmh_simulate_visit_stats": (c_int, [c_void_p, c_void_p, ctypes.POINTER(T.rmh_transform), c_void_p, c_u32,
                                         ctypes.POINTER(ctypes.c_uint64), c_void_p]),
    "rmh_simulate": (c_int, [c_void_p, c_void_p, ctypes.POINTER(T.rmh_transform), c_void_p, c_u32,
                             ctypes.POINTER(T.rmh_outputs), c_void_p]),
    "rmh_batch_create": (c_int, [c_void_p, ctypes.POINTER(T.rmh_batch_entry), c_u32,
                                 ctypes.POINTER(T.rmh_batch_params), ctypes.POINTER(c_void_p)]),
    "rmh_batch_set_params": (c_int, [c_void_p, ctypes.POINTER(T.rmh_batch_params)]),
    "rmh_batch_get_info": (c_int, [c_void_p, ctypes.POINTER(ctypes.c_uint64), ctypes.POINTER(c_u32),
                                   ctypes.POINTER(T.rmh_batch_params)]),
    "rmh_batch_simulate": (c_int, [c_void_p, c_void_p]),
    "rmh_batch_destroy": (c_int, [c_void_p]),
    "rmh_graph_begin_capture": (c_int, [c_void_p]),
    "rmh_graph_end_capture": (c_int, [c_void_p, ctypes.POINTER(c_void_p)]),
    "rmh_graph_launch": (c_int, [c_void_p, c_void_p]),
    "rmh_graph_destroy": (c_int, [c_void_p]),
    "rmh_pose_grad_workspace_size": (c_int, [c_void_p, c_void_p, c_u32, ctypes.POINTER(c_size)]),
    "rmh_pose_grad": (c_int, [c_void_p, c_void_p, ctypes.POINTER(T.rmh_transform), c_void_p, c_u32, c_void_p,
                              c_void_p, c_int, ctypes.c_float, c_void_p, c_void_p, c_size, c_void_p]),
    "rmh_transform_set_identity": (c_int, [c_void_p, c_u32, c_void_p]),
    "rmh_transform_mul": (c_int, [c_void_p, c_u32, c_void_p, c_u32, c_void_p, c_u32, c_void_p]),
    "rmh_transform_inv": (c_int, [c_void_p, c_void_p, c_u32, c_void_p]),
    "rmh_transform_points": (c_int, [c_void_p, c_u32, c_void_p, c_void_p, c_u32, c_void_p]),
    "rmh_quat_normalize": (c_int, [c_void_p, c_u32, c_void_p]),
    "rmh_reduction_workspace_size": (c_int, [ctypes.c_uint64, c_u32, ctypes.POINTER(c_size)]),
    "rmh_vec3_moments": (c_int, [c_void_p, c_void_p, c_void_p, ctypes.c_uint64, c_u32, c_void_p, c_void_p,
                                 c_void_p, c_void_p, c_void_p, c_size, c_void_p]),
    "rmh_statistics": (c_int, [c_int, c_void_p, c_u32, ctypes.POINTER(T.rmh_cloud), ctypes.POINTER(T.rmh_cloud),
                               ctypes.c_uint64, c_u32, ctypes.POINTER(T.rmh_umeyama_constraints), c_void_p,
                               c_void_p, c_size, c_void_p]),
    "rmh_umeyama": (c_int, [c_void_p, c_u32, c_void_p, c_void_p, c_void_p]),
    "rmh_statistics_noalloc": (c_int, [c_int, c_void_p, c_u32, ctypes.POINTER(T.rmh_cloud),
                                       ctypes.POINTER(T.rmh_cloud), ctypes.c_uint64, c_u32,
                                       ctypes.POINTER(T.rmh_umeyama_constraints), c_void_p, c_void_p]),
    "rmh_umeyama_host": (c_int, [c_void_p, c_u32, c_void_p, c_void_p]),
    "rmh_cross_stats_to_f32": (c_int, [c_void_p, c_void_p, c_u32, c_void_p]),
    "rmh_cross_stats_pack": (c_int, [c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_u32, c_void_p]),
    "rmh_transform_pack_matrix": (c_int, [c_void_p, c_void_p, c_void_p, c_u32, c_void_p]),
    "rmh_transform_pack_quat": (c_int, [c_void_p, c_void_p, c_void_p, c_u32, c_void_p]),
    "rmh_statistics_objectwise_p2l": (c_int, [c_void_p, ctypes.POINTER(T.rmh_cloud), ctypes.POINTER(T.rmh_cloud), c_u32,
                                              c_u32, c_u32, c_void_p, c_void_p, c_u32, c_u32, c_void_p, c_void_p]),
    "rmh_linear_system_p2l": (c_int, [c_void_p, c_u32, ctypes.POINTER(T.rmh_cloud), ctypes.POINTER(T.rmh_cloud),
                                      ctypes.c_uint64, c_u32, ctypes.POINTER(T.rmh_umeyama_constraints),
                                      ctypes.c_double, c_void_p, c_void_p, c_size, c_void_p]),
}


def exported_symbols() -> list[str]:
    return list(_SIGNATURES)


def lib() -> ctypes.CDLL:
    """Loads librmh.so once; raises if it has not been built."""
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            raise RuntimeError(f"{LIB_PATH} is missing: build it with `python -m rmagine_amd.build` "
                               "(there is no CPU fallback)")
        _lib = ctypes.CDLL(str(LIB_PATH), mode=os.RTLD_NOW | os.RTLD_GLOBAL)
        for name, (res, args) in _SIGNATURES.items():
            fn = getattr(_lib, name)
            fn.restype = res
            fn.argtypes = args
    return _lib


def check(code: int) -> None:
    if code != 0:
        msg = lib().rmh_last_error().decode(errors="replace")
        if code == 3:
            raise NoMapError(code, msg)
        raise RmhError(code, msg)


def call(name: str, *args) -> None:
    check(getattr(lib(), name)(*args))
