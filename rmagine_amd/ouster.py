"""Ouster sensor metadata -> O1Dn model: the counterpart of o1dn_from_ouster_meta_file
(src/rmagine_ouster/src/types/ouster_sensors.cpp:8-80), over librmh's rmh_ouster_o1dn."""
from __future__ import annotations

import ctypes

import numpy as np

from . import _lib
from . import types as T


def o1dn_from_ouster_meta_file(path) -> T.O1DnModel:
    """O1Dn model of the sensor described by an Ouster metadata JSON file: lidar_mode
    columns x beams, directions from the beam altitude/azimuth tables, origin (0, 0, 0),
    range [0.1, 80] (the reference's values)."""
    w, h = ctypes.c_uint32(), ctypes.c_uint32()
    p = str(path).encode()
    _lib.call("rmh_ouster_o1dn", p, ctypes.byref(w), ctypes.byref(h), None, 0)
    dirs = np.empty((w.value * h.value, 3), np.float32)
    _lib.call("rmh_ouster_o1dn", p, ctypes.byref(w), ctypes.byref(h), ctypes.c_void_p(dirs.ctypes.data), dirs.size)
    return T.O1DnModel(w.value, h.value, T.Interval(0.1, 80.0), (0.0, 0.0, 0.0), dirs)
