"""Python mirror of Rmagine's GPU math layer on librmh.

  statistics_p2p / statistics_p2l   src/rmagine_cuda/include/rmagine/math/statistics.cuh:76-158
  umeyama_transform                 src/rmagine_core/include/rmagine/math/optimization.h:61-75
  multNxN / multNx1 / mult1xN, invert, normalizeInplace, sum / mean / cov / covBatched
                                    src/rmagine_cuda/include/rmagine/math/memory_math.cuh,
                                    math_batched.cuh:54-142

All arrays are torch device tensors.  Transforms are float32 (N, 8); vectors float32
(N, 3); statistics are float64 (N, 16) rows laid out like rmh_cross_stats
(dataset_mean[3], model_mean[3], covariance[9] column-major, n_meas as uint32 in the
low half of the last element).  Segments: element i of segment s is at
s * segment_stride + i of each cloud (stride 0 = shared by every segment).
"""
from __future__ import annotations

import ctypes
from dataclasses import dataclass

import numpy as np
import torch

from . import _lib
from . import types as T
from .simulation import stream_handle

P2P, P2L = 0, 1


def _ptr(t) -> ctypes.c_void_p:
    return ctypes.c_void_p(0 if t is None else int(t.data_ptr()))


def _check_dev(t, dtype, name, cols=None):
    if t is None:
        return
    if not isinstance(t, torch.Tensor) or t.device.type != "cuda" or not t.is_contiguous() or t.dtype != dtype:
        raise ValueError(f"{name} must be a contiguous {dtype} device tensor")
    if cols is not None and (t.dim() != 2 or t.shape[1] != cols):
        raise ValueError(f"{name} must have shape (N, {cols})")


@dataclass
class PointCloudView:
    """PointCloudView_ (src/rmagine_core/include/rmagine/types/PointCloud.hpp:62-70)."""
    points: torch.Tensor
    mask: torch.Tensor | None = None
    normals: torch.Tensor | None = None
    ids: torch.Tensor | None = None
    segment_stride: int | None = None  # None: the segment size

    def to_c(self, segment_size: int) -> T.rmh_cloud:
        _check_dev(self.points, torch.float32, "points", 3)
        _check_dev(self.normals, torch.float32, "normals", 3)
        _check_dev(self.mask, torch.uint8, "mask")
        _check_dev(self.ids, torch.int32, "ids")
        stride = segment_size if self.segment_stride is None else int(self.segment_stride)
        return T.rmh_cloud(int(self.points.data_ptr()), _ptr(self.mask).value or 0, _ptr(self.normals).value or 0,
                           _ptr(self.ids).value or 0, stride)

    def max_index(self, segment_size: int, n_segments: int) -> int:
        stride = segment_size if self.segment_stride is None else int(self.segment_stride)
        return stride * (n_segments - 1) + segment_size


@dataclass
class UmeyamaReductionConstraints:
    """src/rmagine_core/include/rmagine/types/UmeyamaReductionConstraints.hpp:46-59"""
    max_dist: float = float("inf")
    dataset_id: int = 0
    model_id: int = 0


def _workspace(segment_size: int, n_segments: int, device) -> torch.Tensor | None:
    n = ctypes.c_size_t()
    _lib.call("rmh_reduction_workspace_size", segment_size, n_segments, ctypes.byref(n))
    if n.value == 0:
        return None
    return torch.empty(n.value, dtype=torch.uint8, device=device)


def statistics(kind: int, pre_transforms: torch.Tensor, dataset: PointCloudView, model: PointCloudView,
               params: UmeyamaReductionConstraints, segment_size: int | None = None, n_segments: int = 1,
               out: torch.Tensor | None = None, stream=None) -> torch.Tensor:
    """Segmented FP64 p2p/p2l correspondence statistics; returns (n_segments, 16) float64."""
    if segment_size is None:
        segment_size = dataset.points.shape[0] // n_segments
    _check_dev(pre_transforms, torch.float32, "pre_transforms", 8)
    n_pre = pre_transforms.shape[0]
    if n_pre not in (1, n_segments):
        raise ValueError("pre_transforms must hold 1 or n_segments transforms")
    for name, cloud in (("dataset", dataset), ("model", model)):
        need = cloud.max_index(segment_size, n_segments)
        for f in ("points", "mask", "normals", "ids"):
            t = getattr(cloud, f)
            if t is not None and t.shape[0] < need:
                raise ValueError(f"{name}.{f} holds {t.shape[0]} elements, the segments need {need}")
    if kind == P2L and model.normals is None:
        raise ValueError("point-to-plane statistics need model normals")
    dev = dataset.points.device
    if out is None:
        out = torch.empty((n_segments, 16), dtype=torch.float64, device=dev)
    _check_dev(out, torch.float64, "out", 16)
    ws = _workspace(segment_size, n_segments, dev)
    dc, mc = dataset.to_c(segment_size), model.to_c(segment_size)
    prm = T.rmh_umeyama_constraints(params.max_dist, params.dataset_id, params.model_id)
    _lib.call("rmh_statistics", kind, _ptr(pre_transforms), n_pre, ctypes.byref(dc), ctypes.byref(mc),
              segment_size, n_segments, ctypes.byref(prm), _ptr(out), _ptr(ws),
              0 if ws is None else ws.numel(), stream_handle(stream))
    return out


def statistics_p2p(pre_transform, dataset, model, params, **kw):
    return statistics(P2P, pre_transform, dataset, model, params, **kw)


def statistics_p2l(pre_transform, dataset, model, params, **kw):
    return statistics(P2L, pre_transform, dataset, model, params, **kw)


def stats_n_meas(stats: torch.Tensor) -> torch.Tensor:
    """n_meas column of a (N, 16) statistics tensor."""
    return stats.view(torch.int32)[:, 30]


def umeyama_transform(stats: torch.Tensor, out: torch.Tensor | None = None, out64: torch.Tensor | None = None,
                      stream=None):
    """FP64 Umeyama per statistics row; returns (float32 (N,8) transforms, float64 (N,7) or None)."""
    _check_dev(stats, torch.float64, "stats", 16)
    n = stats.shape[0]
    if out is None:
        out = torch.empty((n, 8), dtype=torch.float32, device=stats.device)
    _lib.call("rmh_umeyama", _ptr(stats), n, _ptr(out), _ptr(out64), stream_handle(stream))
    return out, out64


def transform_mul(A: torch.Tensor, B: torch.Tensor, out=None, stream=None) -> torch.Tensor:
    """multNxN / multNx1 / mult1xN of transforms (broadcast a size-1 operand)."""
    _check_dev(A, torch.float32, "A", 8)
    _check_dev(B, torch.float32, "B", 8)
    n = max(A.shape[0], B.shape[0])
    if out is None:
        out = torch.empty((n, 8), dtype=torch.float32, device=A.device)
    _lib.call("rmh_transform_mul", _ptr(A), A.shape[0], _ptr(B), B.shape[0], _ptr(out), n, stream_handle(stream))
    return out


def transform_inv(A: torch.Tensor, out=None, stream=None) -> torch.Tensor:
    _check_dev(A, torch.float32, "A", 8)
    if out is None:
        out = torch.empty_like(A)
    _lib.call("rmh_transform_inv", _ptr(A), _ptr(out), A.shape[0], stream_handle(stream))
    return out


def transform_points(Tr: torch.Tensor, p: torch.Tensor, out=None, stream=None) -> torch.Tensor:
    _check_dev(Tr, torch.float32, "T", 8)
    _check_dev(p, torch.float32, "p", 3)
    if out is None:
        out = torch.empty_like(p)
    _lib.call("rmh_transform_points", _ptr(Tr), Tr.shape[0], _ptr(p), _ptr(out), p.shape[0], stream_handle(stream))
    return out


def quat_normalize_(Tr: torch.Tensor, stream=None) -> torch.Tensor:
    _check_dev(Tr, torch.float32, "T", 8)
    _lib.call("rmh_quat_normalize", _ptr(Tr), Tr.shape[0], stream_handle(stream))
    return Tr


def vec3_moments(a: torch.Tensor, b: torch.Tensor | None = None, mask: torch.Tensor | None = None,
                 segment_size: int | None = None, n_segments: int = 1, stream=None) -> dict:
    """Segmented FP64 sum, mean, cov = (1/n) sum b a^T (column-major 9) and count."""
    _check_dev(a, torch.float32, "a", 3)
    _check_dev(b, torch.float32, "b", 3)
    _check_dev(mask, torch.uint8, "mask")
    if segment_size is None:
        segment_size = a.shape[0] // n_segments
    if segment_size * n_segments > a.shape[0]:
        raise ValueError("segments exceed the input")
    dev = a.device
    res = {"sum": torch.empty((n_segments, 3), dtype=torch.float64, device=dev),
           "mean": torch.empty((n_segments, 3), dtype=torch.float64, device=dev),
           "cov": torch.empty((n_segments, 9), dtype=torch.float64, device=dev),
           "count": torch.empty((n_segments,), dtype=torch.int32, device=dev)}
    ws = _workspace(segment_size, n_segments, dev)
    _lib.call("rmh_vec3_moments", _ptr(a), _ptr(b), _ptr(mask), segment_size, n_segments, _ptr(res["sum"]),
              _ptr(res["mean"]), _ptr(res["cov"]), _ptr(res["count"]), _ptr(ws), 0 if ws is None else ws.numel(),
              stream_handle(stream))
    return res


def sum(a, segment_size=None, n_segments=1, **kw):  # noqa: A001 - mirrors rmagine::sum
    return vec3_moments(a, segment_size=segment_size, n_segments=n_segments, **kw)["sum"]


def mean(a, segment_size=None, n_segments=1, **kw):
    return vec3_moments(a, segment_size=segment_size, n_segments=n_segments, **kw)["mean"]


def cov(v1, v2, mask=None, segment_size=None, n_segments=1, **kw):
    """cov(v1, v2) = (1/N) sum v2 v1^T per segment (memory_math.cu:415-469 layout)."""
    return vec3_moments(v1, v2, mask, segment_size=segment_size, n_segments=n_segments, **kw)["cov"]


def stats_to_numpy(stats: torch.Tensor) -> dict:
    s = stats.detach().cpu().numpy()
    n = s.view(np.uint32)[:, 30].copy()
    return {"dataset_mean": s[:, 0:3], "model_mean": s[:, 3:6],
            "covariance": s[:, 6:15].reshape(-1, 3, 3).transpose(0, 2, 1), "n_meas": n}


def pack(R: torch.Tensor, t: torch.Tensor, out: torch.Tensor | None = None, stream=None) -> torch.Tensor:
    """pack (src/rmagine_cuda/include/rmagine/math/memory_math.cuh:459-470): transforms from
    rotations and translations.  R is (N, 3, 3) float32 rotation matrices (R[i, r, c] = at(r, c),
    converted with the reference's FP32 matrix -> quaternion steps) or (N, 4) quaternions
    (x, y, z, w); t is (N, 3).  `out` (N, 8) keeps its stamps; a new one has stamp 0."""
    _check_dev(t, torch.float32, "t", 3)
    n = t.shape[0]
    if out is None:
        out = torch.zeros((n, 8), dtype=torch.float32, device=t.device)
    _check_dev(out, torch.float32, "out", 8)
    if out.shape[0] != n or R.shape[0] != n:
        raise ValueError("R, t and out must hold the same number of entries")
    if R.dim() == 3 and tuple(R.shape[1:]) == (3, 3):
        if R.dtype != torch.float32 or R.device.type != "cuda":
            raise ValueError("R must be a float32 device tensor")
        Rc = R.transpose(1, 2).contiguous()  # Rmagine's Matrix3x3 is column-major
        _lib.call("rmh_transform_pack_matrix", _ptr(Rc), _ptr(t), _ptr(out), n, stream_handle(stream))
    else:
        _check_dev(R, torch.float32, "R", 4)
        _lib.call("rmh_transform_pack_quat", _ptr(R), _ptr(t), _ptr(out), n, stream_handle(stream))
    return out


def _constraints_array(params, n: int, device) -> torch.Tensor:
    if isinstance(params, UmeyamaReductionConstraints):
        params = [params] * n
    if len(params) != n:
        raise ValueError(f"need {n} constraints, got {len(params)}")
    arr = (T.rmh_umeyama_constraints * n)()
    for i, p in enumerate(params):
        arr[i] = T.rmh_umeyama_constraints(p.max_dist, p.dataset_id, p.model_id)
    host = torch.frombuffer(bytearray(bytes(arr)), dtype=torch.uint8)
    return host.to(device)


def statistics_objectwise_p2l(pre_transforms: torch.Tensor, dataset: PointCloudView, model: PointCloudView,
                              width: int, height: int, params, bboxes: torch.Tensor, window_only: bool = False,
                              out: torch.Tensor | None = None, stream=None) -> torch.Tensor:
    """statistics_objectwise_p2l (src/rmagine_cuda/include/rmagine/math/statistics.cuh:139-158):
    one FP64 statistics row per object over its pixel window (bboxes (n, 6) float32: min xyz,
    max xyz; x columns, y rows) of image-shaped scans with row stride `width`, with per-object
    pre-transforms (n, 8) and constraints (one UmeyamaReductionConstraints or a list of n).
    Returns (n, 16) float64."""
    _check_dev(pre_transforms, torch.float32, "pre_transforms", 8)
    _check_dev(bboxes, torch.float32, "bboxes", 6)
    n = pre_transforms.shape[0]
    if bboxes.shape[0] != n:
        raise ValueError("one bounding box per pre-transform")
    if model.normals is None:
        raise ValueError("point-to-plane statistics need model normals")
    n_points = dataset.points.shape[0]
    for name, cloud in (("dataset", dataset), ("model", model)):
        for f in ("points", "mask", "normals", "ids"):
            tt = getattr(cloud, f)
            if tt is not None and tt.shape[0] < n_points:
                raise ValueError(f"{name}.{f} holds fewer elements than dataset.points")
    dev = pre_transforms.device
    prm = _constraints_array(params, n, dev)
    if out is None:
        out = torch.empty((n, 16), dtype=torch.float64, device=dev)
    _check_dev(out, torch.float64, "out", 16)
    dc, mc = dataset.to_c(0), model.to_c(0)
    _lib.call("rmh_statistics_objectwise_p2l", _ptr(pre_transforms), ctypes.byref(dc), ctypes.byref(mc), int(width),
              int(height), int(n_points), _ptr(prm), _ptr(bboxes), n, 1 if window_only else 0, _ptr(out),
              stream_handle(stream))
    return out


P2L_SYSTEM_COLS = 44  # rmh_p2l_system: JTwJ (36), JTwr (6), wr2, n_meas (uint32) + pad


def linear_system_p2l(pre_transforms: torch.Tensor, dataset: PointCloudView, model: PointCloudView,
                      params: UmeyamaReductionConstraints, kernel_scale: float = 5.0, segment_size: int | None = None,
                      n_segments: int = 1, out: torch.Tensor | None = None, workspace: torch.Tensor | None = None,
                      stream=None) -> torch.Tensor:
    """Segmented FP64 point-to-plane Gauss-Newton system (build_linear_system_p2l,
    src/rmagine_core/include/rmagine/math/optimization.h:170-200).  Returns (n_segments, 44)
    float64 rows laid out like rmh_p2l_system (JTwJ row-major 6x6, JTwr, sum w r^2, n_meas)."""
    if segment_size is None:
        segment_size = dataset.points.shape[0] // n_segments
    _check_dev(pre_transforms, torch.float32, "pre_transforms", 8)
    n_pre = pre_transforms.shape[0]
    if n_pre not in (1, n_segments):
        raise ValueError("pre_transforms must hold 1 or n_segments transforms")
    if model.normals is None:
        raise ValueError("point-to-plane system needs model normals")
    for name, cloud in (("dataset", dataset), ("model", model)):
        need = cloud.max_index(segment_size, n_segments)
        for f in ("points", "mask", "normals", "ids"):
            tt = getattr(cloud, f)
            if tt is not None and tt.shape[0] < need:
                raise ValueError(f"{name}.{f} holds {tt.shape[0]} elements, the segments need {need}")
    dev = dataset.points.device
    if out is None:
        out = torch.empty((n_segments, P2L_SYSTEM_COLS), dtype=torch.float64, device=dev)
    _check_dev(out, torch.float64, "out", P2L_SYSTEM_COLS)
    ws = workspace if workspace is not None else _workspace(segment_size, n_segments, dev)
    dc, mc = dataset.to_c(segment_size), model.to_c(segment_size)
    prm = T.rmh_umeyama_constraints(params.max_dist, params.dataset_id, params.model_id)
    _lib.call("rmh_linear_system_p2l", _ptr(pre_transforms), n_pre, ctypes.byref(dc), ctypes.byref(mc), segment_size,
              n_segments, ctypes.byref(prm), float(kernel_scale), _ptr(out), _ptr(ws),
              0 if ws is None else ws.numel(), stream_handle(stream))
    return out


def p2l_system_to_numpy(sysm: torch.Tensor) -> dict:
    s = sysm.detach().cpu().numpy()
    return {"JTwJ": s[:, 0:36].reshape(-1, 6, 6), "JTwr": s[:, 36:42], "wr2": s[:, 42],
            "n_meas": s.view(np.uint32)[:, 86].copy()}
