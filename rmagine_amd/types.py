"""Sensor models, transforms and result bundles, mirroring Rmagine's types.

Field names and meanings follow the reference:
  Interval, DiscreteInterval, SphericalModel, PinholeModel, O1DnModel_, OnDnModel_
      src/rmagine_core/include/rmagine/types/sensor_models.h:55-503
  Transform {R (quaternion x,y,z,w), t, stamp}, 32 bytes
      src/rmagine_core/include/rmagine/math/types/Transform.hpp:63-68
  Hits/Ranges/Points/Normals/FaceIds/GeomIds/ObjectIds and IntAttrAll
      src/rmagine_core/include/rmagine/simulation/SimulationResults.hpp:55-163
  example_spherical/pinhole/o1dn/ondn, vlp16_900
      src/rmagine_core/src/types/sensors.cpp:6-137

All floating point fields are stored as FP32 values (what the reference's structs hold).
Transforms are float32 arrays of shape (N, 8): qx, qy, qz, qw, tx, ty, tz, stamp.
"""
from __future__ import annotations

import ctypes
import math
from dataclasses import dataclass, field

import numpy as np

DEG_TO_RAD = math.pi / 180.0
DEG_TO_RAD_F = float(np.float32(DEG_TO_RAD))

MODEL_SPHERICAL, MODEL_PINHOLE, MODEL_O1DN, MODEL_ONDN = 0, 1, 2, 3

ATTRIBUTES = ("hits", "ranges", "points", "normals", "face_ids", "geom_ids", "object_ids")
ATTR_DTYPES = {
    "hits": (np.uint8, 1),
    "ranges": (np.float32, 1),
    "points": (np.float32, 3),
    "normals": (np.float32, 3),
    "face_ids": (np.uint32, 1),
    "geom_ids": (np.uint32, 1),
    "object_ids": (np.uint32, 1),
}
UINT_MAX = 0xFFFFFFFF


def f32(x: float) -> float:
    return float(np.float32(x))


@dataclass
class Interval:
    min: float
    max: float

    def invalidValue(self) -> float:  # sensor_models.h:59-63
        return f32(np.float64(np.float32(self.max)) + 1.0)


@dataclass
class DiscreteInterval:
    min: float
    inc: float
    size: int

    def getValue(self, i: int) -> float:  # sensor_models.h:89-92, FP32 arithmetic
        return float(np.float32(self.min) + np.float32(i) * np.float32(self.inc))


@dataclass
class SphericalModel:
    phi: DiscreteInterval
    theta: DiscreteInterval
    range: Interval
    kind: int = field(default=MODEL_SPHERICAL, init=False)

    def getWidth(self) -> int:
        return int(self.theta.size)

    def getHeight(self) -> int:
        return int(self.phi.size)

    def size(self) -> int:
        return self.getWidth() * self.getHeight()

    def getBufferId(self, vid: int, hid: int) -> int:
        return vid * self.getWidth() + hid


@dataclass
class PinholeModel:
    width: int
    height: int
    range: Interval
    f: tuple
    c: tuple
    kind: int = field(default=MODEL_PINHOLE, init=False)

    def getWidth(self) -> int:
        return int(self.width)

    def getHeight(self) -> int:
        return int(self.height)

    def size(self) -> int:
        return self.getWidth() * self.getHeight()

    def getBufferId(self, vid: int, hid: int) -> int:
        return vid * self.getWidth() + hid


@dataclass
class O1DnModel:
    width: int
    height: int
    range: Interval
    orig: tuple
    dirs: np.ndarray  # (width*height, 3) float32
    kind: int = field(default=MODEL_O1DN, init=False)

    def getWidth(self) -> int:
        return int(self.width)

    def getHeight(self) -> int:
        return int(self.height)

    def size(self) -> int:
        return self.getWidth() * self.getHeight()

    def getBufferId(self, vid: int, hid: int) -> int:
        return vid * self.getWidth() + hid


@dataclass
class OnDnModel:
    width: int
    height: int
    range: Interval
    origs: np.ndarray  # (width*height, 3) float32
    dirs: np.ndarray   # (width*height, 3) float32
    kind: int = field(default=MODEL_ONDN, init=False)

    def getWidth(self) -> int:
        return int(self.width)

    def getHeight(self) -> int:
        return int(self.height)

    def size(self) -> int:
        return self.getWidth() * self.getHeight()

    def getBufferId(self, vid: int, hid: int) -> int:
        return vid * self.getWidth() + hid


# ---- ctypes mirrors of include/rmh.h -------------------------------------------------
class rmh_vec3(ctypes.Structure):
    _fields_ = [("x", ctypes.c_float), ("y", ctypes.c_float), ("z", ctypes.c_float)]


class rmh_quat(ctypes.Structure):
    _fields_ = [("x", ctypes.c_float), ("y", ctypes.c_float), ("z", ctypes.c_float), ("w", ctypes.c_float)]


class rmh_transform(ctypes.Structure):
    _fields_ = [("R", rmh_quat), ("t", rmh_vec3), ("stamp", ctypes.c_uint32)]


class rmh_interval(ctypes.Structure):
    _fields_ = [("min", ctypes.c_float), ("max", ctypes.c_float)]


class rmh_discrete_interval(ctypes.Structure):
    _fields_ = [("min", ctypes.c_float), ("inc", ctypes.c_float), ("size", ctypes.c_uint32)]


class rmh_spherical_model(ctypes.Structure):
    _fields_ = [("phi", rmh_discrete_interval), ("theta", rmh_discrete_interval), ("range", rmh_interval)]


class rmh_pinhole_model(ctypes.Structure):
    _fields_ = [("width", ctypes.c_uint32), ("height", ctypes.c_uint32), ("range", rmh_interval),
                ("f", ctypes.c_float * 2), ("c", ctypes.c_float * 2)]


class rmh_o1dn_model(ctypes.Structure):
    _fields_ = [("width", ctypes.c_uint32), ("height", ctypes.c_uint32), ("range", rmh_interval),
                ("orig", rmh_vec3), ("_pad", ctypes.c_uint32), ("dirs_dev", ctypes.c_void_p)]


class rmh_ondn_model(ctypes.Structure):
    _fields_ = [("width", ctypes.c_uint32), ("height", ctypes.c_uint32), ("range", rmh_interval),
                ("origs_dev", ctypes.c_void_p), ("dirs_dev", ctypes.c_void_p)]


class rmh_outputs(ctypes.Structure):
    _fields_ = [(name, ctypes.c_void_p) for name in ATTRIBUTES]


BVH_WIDE = 1  # rmh_bvh_params.flags: also build the 4-wide quantised tree and traverse it
BVH_NO_REFINE = 2  # rmh_bvh_params.flags: skip the treelet SAH refinement of the radix tree


class rmh_bvh_params(ctypes.Structure):
    _fields_ = [("max_leaf_size", ctypes.c_uint32), ("sah_cost_node", ctypes.c_float),
                ("sah_cost_tri", ctypes.c_float), ("flags", ctypes.c_uint32)]


class rmh_map_info(ctypes.Structure):
    _fields_ = [("n_vertices", ctypes.c_uint32), ("n_faces", ctypes.c_uint32), ("n_meshes", ctypes.c_uint32),
                ("n_nodes", ctypes.c_uint32), ("n_leaves", ctypes.c_uint32), ("max_depth", ctypes.c_uint32),
                ("device_bytes", ctypes.c_uint64), ("build_ms", ctypes.c_float), ("sah_cost", ctypes.c_float),
                ("n_wide_nodes", ctypes.c_uint32), ("wide_levels", ctypes.c_uint32)]


class rmh_cloud(ctypes.Structure):
    _fields_ = [("points", ctypes.c_void_p), ("mask", ctypes.c_void_p), ("normals", ctypes.c_void_p),
                ("ids", ctypes.c_void_p), ("segment_stride", ctypes.c_uint64)]


class rmh_umeyama_constraints(ctypes.Structure):
    _fields_ = [("max_dist", ctypes.c_float), ("dataset_id", ctypes.c_uint32), ("model_id", ctypes.c_uint32)]


class rmh_cross_stats(ctypes.Structure):
    _fields_ = [("dataset_mean", ctypes.c_double * 3), ("model_mean", ctypes.c_double * 3),
                ("covariance", ctypes.c_double * 9), ("n_meas", ctypes.c_uint32), ("_pad", ctypes.c_uint32)]


class rmh_aabb(ctypes.Structure):
    _fields_ = [("min", rmh_vec3), ("max", rmh_vec3)]


class rmh_p2l_system(ctypes.Structure):
    _fields_ = [("JTwJ", ctypes.c_double * 36), ("JTwr", ctypes.c_double * 6), ("wr2", ctypes.c_double),
                ("n_meas", ctypes.c_uint32), ("_pad", ctypes.c_uint32)]


class rmh_batch_entry(ctypes.Structure):
    _fields_ = [("sensor", ctypes.c_void_p), ("Tsb", rmh_transform), ("Tbm_dev", ctypes.c_void_p),
                ("n_poses", ctypes.c_uint32), ("_pad", ctypes.c_uint32), ("out", rmh_outputs)]


class rmh_batch_params(ctypes.Structure):
    _fields_ = [("chunk_rays", ctypes.c_uint32), ("waves_per_cu", ctypes.c_uint32), ("tile_rows", ctypes.c_uint32),
                ("lds_stack", ctypes.c_uint32), ("launches", ctypes.c_uint32)]


assert ctypes.sizeof(rmh_cross_stats) == 128
assert ctypes.sizeof(rmh_transform) == 32
assert ctypes.sizeof(rmh_spherical_model) == 32
assert ctypes.sizeof(rmh_pinhole_model) == 32


def _interval(r: Interval) -> rmh_interval:
    return rmh_interval(r.min, r.max)


def model_to_c(model, dirs_dev: int = 0, origs_dev: int = 0):
    """Returns (kind, ctypes struct) for a model; device pointers for O1Dn/OnDn rays."""
    if isinstance(model, SphericalModel):
        s = rmh_spherical_model(rmh_discrete_interval(model.phi.min, model.phi.inc, model.phi.size),
                                rmh_discrete_interval(model.theta.min, model.theta.inc, model.theta.size),
                                _interval(model.range))
        return MODEL_SPHERICAL, s
    if isinstance(model, PinholeModel):
        s = rmh_pinhole_model(model.width, model.height, _interval(model.range),
                              (ctypes.c_float * 2)(*model.f), (ctypes.c_float * 2)(*model.c))
        return MODEL_PINHOLE, s
    if isinstance(model, O1DnModel):
        s = rmh_o1dn_model(model.width, model.height, _interval(model.range), rmh_vec3(*model.orig), 0, dirs_dev)
        return MODEL_O1DN, s
    if isinstance(model, OnDnModel):
        s = rmh_ondn_model(model.width, model.height, _interval(model.range), origs_dev, dirs_dev)
        return MODEL_ONDN, s
    raise TypeError(f"unsupported sensor model {type(model)!r}")


def model_to_floats(model) -> np.ndarray:
    """Flat FP32 parameter vector (the oracle's calling convention)."""
    if isinstance(model, SphericalModel):
        v = [model.phi.min, model.phi.inc, model.phi.size, model.theta.min, model.theta.inc, model.theta.size,
             model.range.min, model.range.max]
    elif isinstance(model, PinholeModel):
        v = [model.width, model.height, model.range.min, model.range.max, model.f[0], model.f[1], model.c[0],
             model.c[1]]
    elif isinstance(model, O1DnModel):
        v = [model.width, model.height, model.range.min, model.range.max, *model.orig, 0.0]
    elif isinstance(model, OnDnModel):
        v = [model.width, model.height, model.range.min, model.range.max, 0.0, 0.0, 0.0, 0.0]
    else:
        raise TypeError(type(model))
    return np.asarray(v, dtype=np.float32)


# ---- example models (src/rmagine_core/src/types/sensors.cpp) ---------------------------
def vlp16_900() -> SphericalModel:  # sensors.cpp:6-20
    return SphericalModel(
        phi=DiscreteInterval(f32(-15.0 * DEG_TO_RAD_F), f32(2.0 * DEG_TO_RAD_F), 16),
        theta=DiscreteInterval(f32(-math.pi), f32((360.0 / 900.0) * DEG_TO_RAD_F), 900),
        range=Interval(0.5, 130.0))


def example_spherical() -> SphericalModel:  # sensors.cpp:39-54
    return SphericalModel(
        phi=DiscreteInterval(f32(-15.0 * math.pi / 180.0), f32(2.0 * math.pi / 180.0), 16),
        theta=DiscreteInterval(f32(-math.pi), f32(0.4 * math.pi / 180.0), 900),
        range=Interval(0.0, 100.0))


def example_pinhole() -> PinholeModel:  # sensors.cpp:56-68
    return PinholeModel(width=200, height=150, range=Interval(0.0, 100.0), f=(100.0, 100.0), c=(100.0, 75.0))


def example_o1dn() -> O1DnModel:  # sensors.cpp:70-102 (FP32 arithmetic as in the reference)
    dirs = np.zeros((200, 3), dtype=np.float32)
    step = np.float32(0.05)
    for i in range(200):
        y = np.float32(-np.float32(i - 100) * step)
        x = np.float32(np.float32(math.cos(float(y))) * np.float32(2.0) + np.float32(2.0))
        z = np.float32(-1.0)
        n = np.float32(math.sqrt(float(np.float32(np.float32(x * x) + np.float32(y * y)) + np.float32(z * z))))
        dirs[i] = (x / n, y / n, z / n)
    return O1DnModel(width=200, height=1, range=Interval(0.0, 100.0), orig=(0.0, 0.0, 0.0), dirs=dirs)


def example_ondn() -> OnDnModel:  # sensors.cpp:104-137
    origs = np.zeros((200, 3), dtype=np.float32)
    dirs = np.zeros((200, 3), dtype=np.float32)
    step = np.float32(0.05)
    for i in range(200):
        percent = np.float32(np.float32(i) / np.float32(200))
        st = np.float32(-np.float32(i - 100) * step)
        origs[i] = (0.0, np.float32(math.sin(float(st))) * percent, np.float32(math.cos(float(st))) * percent)
        dirs[i] = (1.0, 0.0, 0.0)
    return OnDnModel(width=200, height=1, range=Interval(0.0, 100.0), origs=origs, dirs=dirs)


# ---- transforms -------------------------------------------------------------------------
def identity_transforms(n: int) -> np.ndarray:
    T = np.zeros((n, 8), dtype=np.float32)
    T[:, 3] = 1.0
    return T


def quat_from_euler(roll: float, pitch: float, yaw: float) -> np.ndarray:
    """EulerAngles -> Quaternion, FP32 (src/rmagine_core/include/rmagine/math/types/EulerAngles.tcc:40-61)."""
    r, p, y = np.float32(roll), np.float32(pitch), np.float32(yaw)
    h = np.float32(2.0)
    cr, sr = np.float32(math.cos(float(r / h))), np.float32(math.sin(float(r / h)))
    cp, sp = np.float32(math.cos(float(p / h))), np.float32(math.sin(float(p / h)))
    cy, sy = np.float32(math.cos(float(y / h))), np.float32(math.sin(float(y / h)))
    return np.array([sr * cp * cy - cr * sp * sy,
                     cr * sp * cy + sr * cp * sy,
                     cr * cp * sy - sr * sp * cy,
                     cr * cp * cy + sr * sp * sy], dtype=np.float32)


def make_transforms(t: np.ndarray, q: np.ndarray) -> np.ndarray:
    t = np.asarray(t, dtype=np.float32).reshape(-1, 3)
    q = np.asarray(q, dtype=np.float32).reshape(-1, 4)
    T = np.zeros((max(len(t), len(q)), 8), dtype=np.float32)
    T[:, 0:4] = q
    T[:, 4:7] = t
    return T


def transform_to_c(T) -> rmh_transform:
    T = np.asarray(T, dtype=np.float32).reshape(8)
    return rmh_transform(rmh_quat(*map(float, T[0:4])), rmh_vec3(*map(float, T[4:7])), 0)


NOISE_GAUSSIAN, NOISE_REL_GAUSSIAN, NOISE_UNIFORM_DUST = 0, 1, 2


class rmh_noise_params(ctypes.Structure):
    _fields_ = [("kind", ctypes.c_uint32), ("p0", ctypes.c_float), ("p1", ctypes.c_float), ("p2", ctypes.c_float),
                ("max_range", ctypes.c_float), ("seed", ctypes.c_uint32)]


class rmh_closest_outputs(ctypes.Structure):
    _fields_ = [("dist", ctypes.c_void_p), ("points", ctypes.c_void_p), ("normals", ctypes.c_void_p),
                ("face_ids", ctypes.c_void_p), ("geom_ids", ctypes.c_void_p), ("object_ids", ctypes.c_void_p)]
