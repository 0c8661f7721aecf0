"""ORACLE -- test infrastructure, not product code.

FP64 NumPy restatement of Rmagine's correspondence statistics and Umeyama solve:
  CrossStatistics merge      src/rmagine_core/include/rmagine/math/types/CrossStatistics.tcc:9-58
  statistics_p2p / _p2l      src/rmagine_core/src/math/statistics.cpp:20-126
                             (GPU twin src/rmagine_cuda/src/math/statistics.cu:58-276)
  umeyama_transform          src/rmagine_core/src/math/optimization.cpp:12-55
  matrix -> quaternion       src/rmagine_core/include/rmagine/math/types/MatrixOps.tcc:1492-1529
  sum / mean / cov           src/rmagine_cuda/src/math/memory_math.cu:415-469,1403-1549
  statistics_objectwise_p2l  src/rmagine_cuda/src/math/statistics.cu:278-371 (element set; then p2l as above)
  build_linear_system_p2l    src/rmagine_core/include/rmagine/math/optimization.h:113-137,170-200 (FP64)
  pack (matrix -> quaternion in FP32)
                             src/rmagine_cuda/src/math/memory_math.cu:313-339, MatrixOps.tcc:1492-1529
The statistics are evaluated in two passes (means, then centred products), a different
summation than the device's shifted raw moments.  The SVD is numpy.linalg.svd (LAPACK).

Pins (tests/test_stats_oracle.py): the reference's paper example (tests/core/math_statistics.cpp:323-380),
the mask/id filter counts n, n/2, 0 (:479-660), merge == direct formula, exact recovery of
known rigid transforms, R^T R = I and det R = +1, n_meas = 0 -> identity.  pack_matrix_f32:
scipy.spatial.transform.Rotation.from_matrix on random and branch-selecting rotations (up to
sign, FP32 rounding).  objectwise_elements: a literal enumeration of the reference kernel's 512
threads on small windows, and window_only == the plain p2l statistics of the window's gathered
elements.  linear_system_p2l: central finite differences of (1/2) sum w r^2 (w held fixed)
under left perturbations (gradient) and of the residual vector (Jacobian, J^T W J).
"""
from __future__ import annotations

import numpy as np


def quat_matrix(q) -> np.ndarray:
    """Matrix of p -> q p q^-1 (the reference's sandwich, Quaternion.tcc:70-75) in FP64."""
    x, y, z, w = (float(v) for v in q)
    return np.array([[w * w + x * x - y * y - z * z, 2 * (x * y - w * z), 2 * (x * z + w * y)],
                     [2 * (x * y + w * z), w * w - x * x + y * y - z * z, 2 * (y * z - w * x)],
                     [2 * (x * z - w * y), 2 * (y * z + w * x), w * w - x * x - y * y + z * z]])


def identity_stats():
    return {"dataset_mean": np.zeros(3), "model_mean": np.zeros(3), "covariance": np.zeros((3, 3)), "n_meas": 0}


def init_stats(d, m):
    return {"dataset_mean": np.asarray(d, np.float64), "model_mean": np.asarray(m, np.float64),
            "covariance": np.zeros((3, 3)), "n_meas": 1}


def merge(a, b):
    """CrossStatistics_::add (CrossStatistics.tcc:9-31)."""
    n = a["n_meas"] + b["n_meas"]
    if n == 0:
        return dict(a)
    w1, w2 = a["n_meas"] / n, b["n_meas"] / n
    dm = a["dataset_mean"] * w1 + b["dataset_mean"] * w2
    mm = a["model_mean"] * w1 + b["model_mean"] * w2
    P1 = np.outer(a["model_mean"] - mm, a["dataset_mean"] - dm)
    P2 = np.outer(b["model_mean"] - mm, b["dataset_mean"] - dm)
    C = (a["covariance"] + P1) * w1 + (b["covariance"] + P2) * w2
    return {"dataset_mean": dm, "model_mean": mm, "covariance": C, "n_meas": n}


def _direct(D, M):
    n = len(D)
    if n == 0:
        return identity_stats()
    dm, mm = D.mean(0), M.mean(0)
    C = (M - mm).T @ (D - dm) / n
    return {"dataset_mean": dm, "model_mean": mm, "covariance": C, "n_meas": n}


def statistics(kind: str, pre, dataset: dict, model: dict, segment_size: int, n_segments: int, max_dist: float,
               dataset_id: int = 0, model_id: int = 0) -> list:
    """kind 'p2p' or 'p2l'; dataset/model: dict(points, mask=None, normals=None, ids=None, stride=None)."""
    pre = np.asarray(pre, np.float32).reshape(-1, 8)
    out = []
    for s in range(n_segments):
        P = pre[0 if len(pre) == 1 else s]
        R = quat_matrix(P[0:4])
        t = P[4:7].astype(np.float64)

        def seg(c, key):
            a = c.get(key)
            if a is None:
                return None
            st = segment_size if c.get("stride") is None else c["stride"]
            return np.asarray(a)[s * st: s * st + segment_size]

        ok = np.ones(segment_size, bool)
        for c, idk in ((dataset, dataset_id), (model, model_id)):
            mk = seg(c, "mask")
            if mk is not None:
                ok &= mk > 0
            ids = seg(c, "ids")
            if ids is not None:
                ok &= ids.astype(np.uint32) == np.uint32(idk)
        D = seg(dataset, "points")[ok].astype(np.float64) @ R.T + t
        I = seg(model, "points")[ok].astype(np.float64)
        if kind == "p2l":
            N = seg(model, "normals")[ok].astype(np.float64)
            spd = np.einsum("ij,ij->i", I - D, N)
            keep = np.abs(spd) < max_dist
            M = D[keep] + N[keep] * spd[keep, None]
            D = D[keep]
        else:
            keep = np.linalg.norm(I - D, axis=1) < max_dist
            D, M = D[keep], I[keep]
        out.append(_direct(D, M))
    return out


def matrix_to_quat(Rm) -> np.ndarray:
    """MatrixOps.tcc:1492-1529 in FP64, returns (x, y, z, w)."""
    R = np.asarray(Rm, np.float64)
    tr = R[0, 0] + R[1, 1] + R[2, 2]
    if tr > 0:
        S = np.sqrt(tr + 1.0) * 2
        w, x, y, z = 0.25 * S, (R[2, 1] - R[1, 2]) / S, (R[0, 2] - R[2, 0]) / S, (R[1, 0] - R[0, 1]) / S
    elif R[0, 0] > R[1, 1] and R[0, 0] > R[2, 2]:
        S = np.sqrt(1.0 + R[0, 0] - R[1, 1] - R[2, 2]) * 2
        w, x, y, z = (R[2, 1] - R[1, 2]) / S, 0.25 * S, (R[0, 1] + R[1, 0]) / S, (R[0, 2] + R[2, 0]) / S
    elif R[1, 1] > R[2, 2]:
        S = np.sqrt(1.0 + R[1, 1] - R[0, 0] - R[2, 2]) * 2
        w, x, y, z = (R[0, 2] - R[2, 0]) / S, (R[0, 1] + R[1, 0]) / S, 0.25 * S, (R[1, 2] + R[2, 1]) / S
    else:
        S = np.sqrt(1.0 + R[2, 2] - R[0, 0] - R[1, 1]) * 2
        w, x, y, z = (R[1, 0] - R[0, 1]) / S, (R[0, 2] + R[2, 0]) / S, (R[1, 2] + R[2, 1]) / S, 0.25 * S
    q = np.array([x, y, z, w])
    return q / np.linalg.norm(q)


def umeyama(st) -> np.ndarray:
    """optimization.cpp:12-49 in FP64; returns (qx, qy, qz, qw, tx, ty, tz)."""
    if st["n_meas"] == 0:
        return np.array([0, 0, 0, 1, 0, 0, 0], np.float64)
    C = np.asarray(st["covariance"], np.float64)
    U, S, Vt = np.linalg.svd(C)
    q = np.array([0, 0, 0, 1.0])
    if S[0] > 0:
        Sd = np.eye(3)
        if np.linalg.det(U) * np.linalg.det(Vt.T) < 0:
            Sd[2, 2] = -1
        q2 = matrix_to_quat(U @ Sd @ Vt)
        if np.all(np.isfinite(q2)):
            q = q2
    t = st["model_mean"] - quat_matrix(q) @ st["dataset_mean"]
    return np.concatenate([q, t])


def moments(a, b=None, mask=None, segment_size=None, n_segments=1):
    a = np.asarray(a, np.float64)
    b = a if b is None else np.asarray(b, np.float64)
    seg = segment_size or len(a) // n_segments
    res = {"sum": [], "mean": [], "cov": [], "count": []}
    for s in range(n_segments):
        sl = slice(s * seg, (s + 1) * seg)
        ok = np.ones(seg, bool) if mask is None else (np.asarray(mask)[sl] > 0)
        A, B = a[sl][ok], b[sl][ok]
        n = len(A)
        res["sum"].append(A.sum(0))
        res["mean"].append(A.mean(0) if n else np.zeros(3))
        C = (B.T @ A) / n if n else np.zeros((3, 3))  # (1/n) sum b a^T
        res["cov"].append(C.T.reshape(9))                # column-major
        res["count"].append(n)
    return {k: np.asarray(v) for k, v in res.items()}


def pack_matrix_f32(Rm) -> np.ndarray:
    """Matrix3x3 -> Quaternion in FP32 (MatrixOps.tcc:1492-1529), as pack_kernel uses it
    (src/rmagine_cuda/src/math/memory_math.cu:313-325).  Rm: (N, 3, 3) float32, [i, r, c] =
    at(r, c).  Every step is one FP32 (or, where the reference writes a double literal, one
    FP64-then-FP32) operation; returns (N, 4) float32 (x, y, z, w), not normalised."""
    Rm = np.asarray(Rm, np.float32).reshape(-1, 3, 3)
    f, d = np.float32, np.float64
    out = np.empty((len(Rm), 4), np.float32)
    for i, M in enumerate(Rm):
        a = [[f(M[r, c]) for c in range(3)] for r in range(3)]
        tr = f(0.0)
        for k in range(3):
            tr = f(tr + a[k][k])
        if tr > 0:
            S = f(np.sqrt(f(d(tr) + 1.0))) * f(2)
            w, x, y, z = f(0.25) * S, (a[2][1] - a[1][2]) / S, (a[0][2] - a[2][0]) / S, (a[1][0] - a[0][1]) / S
        elif a[0][0] > a[1][1] and a[0][0] > a[2][2]:
            S = f(np.sqrt(f(1.0 + d(a[0][0]) - d(a[1][1]) - d(a[2][2])))) * f(2)
            w, x, y, z = (a[2][1] - a[1][2]) / S, f(0.25) * S, (a[0][1] + a[1][0]) / S, (a[0][2] + a[2][0]) / S
        elif a[1][1] > a[2][2]:
            S = f(np.sqrt(f(1.0 + d(a[1][1]) - d(a[0][0]) - d(a[2][2])))) * f(2)
            w, x, y, z = (a[0][2] - a[2][0]) / S, (a[0][1] + a[1][0]) / S, f(0.25) * S, (a[1][2] + a[2][1]) / S
        else:
            S = f(np.sqrt(f(1.0 + d(a[2][2]) - d(a[0][0]) - d(a[1][1])))) * f(2)
            w, x, y, z = (a[1][0] - a[0][1]) / S, (a[0][2] + a[2][0]) / S, (a[1][2] + a[2][1]) / S, f(0.25) * S
        out[i] = (f(x), f(y), f(z), f(w))
    return out


def _f2u_sat(v: float) -> int:
    """float -> unsigned as the GPU's saturating conversion gives it (NaN, <= 0 -> 0)."""
    v = float(np.float32(v))
    if not v > 0.0:
        return 0
    if v >= 4294967296.0:
        return 0xFFFFFFFF
    return int(v)


def objectwise_elements(bbox, width: int, n_points: int, window_only: bool = False) -> np.ndarray:
    """Scan indices statistics_objectwise_p2l_kernel (src/rmagine_cuda/src/math/statistics.cu:
    278-371) visits for one object: window element k -> (min_row + k // w) * width + min_col +
    k % w, k < ceil(w*h / 512) * 512 (its 512 threads take ceil(w*h/512) consecutive k each;
    w*h with window_only), kept while below n_points.  bbox = (min x, y, z, max x, y, z)."""
    c0, r0 = _f2u_sat(bbox[0]), _f2u_sat(bbox[1])
    c1, r1 = _f2u_sat(bbox[3]), _f2u_sat(bbox[4])
    bw, bh = max(c1 - c0, 0), max(r1 - r0, 0)
    n = bw * bh
    if not window_only:
        n = (n + 511) // 512 * 512
    if n == 0:
        return np.zeros(0, np.int64)
    k = np.arange(n, dtype=np.int64)
    ids = (r0 + k // bw) * width + c0 + k % bw
    return ids[ids < n_points]


def statistics_objectwise_p2l(pre, dataset: dict, model: dict, width: int, n_points: int, params, bboxes,
                              window_only: bool = False) -> list:
    """Per object o: statistics_p2l over objectwise_elements(bboxes[o]) with pre[o] and
    params[o] = (max_dist, dataset_id, model_id)."""
    pre = np.asarray(pre, np.float32).reshape(-1, 8)
    out = []
    for o in range(len(pre)):
        idx = objectwise_elements(np.asarray(bboxes[o], np.float32), width, n_points, window_only)

        def take(c):
            return {k: (None if c.get(k) is None else np.asarray(c[k])[idx]) for k in ("points", "mask", "normals", "ids")}

        md, did, mid = params[o]
        out.extend(statistics("p2l", pre[o], take(dataset), take(model), len(idx), 1, float(np.float32(md)),
                              int(did), int(mid)))
    return out


def linear_system_p2l(pre, dataset: dict, model: dict, segment_size: int, n_segments: int, max_dist: float,
                      kernel_scale: float = 5.0, dataset_id: int = 0, model_id: int = 0) -> list:
    """build_linear_system_p2l (src/rmagine_core/include/rmagine/math/optimization.h:170-200)
    with jacobian_and_residual_p2l (:113-137) in FP64 per segment, after the mask/id filters
    and |r| < max_dist: r = (Pd - Pm) . Nm, J = [Pd x Nm, Nm], w = s^2 / (s + r)^2 (s <= 0:
    w = 1).  Returns dicts JTwJ (6, 6), JTwr (6,), wr2, n_meas."""
    pre = np.asarray(pre, np.float32).reshape(-1, 8)
    out = []
    for s in range(n_segments):
        P = pre[0 if len(pre) == 1 else s]
        R = quat_matrix(P[0:4])
        t = P[4:7].astype(np.float64)

        def seg(c, key):
            a = c.get(key)
            if a is None:
                return None
            st = segment_size if c.get("stride") is None else c["stride"]
            return np.asarray(a)[s * st: s * st + segment_size]

        ok = np.ones(segment_size, bool)
        for c, idk in ((dataset, dataset_id), (model, model_id)):
            mk = seg(c, "mask")
            if mk is not None:
                ok &= mk > 0
            ids = seg(c, "ids")
            if ids is not None:
                ok &= ids.astype(np.uint32) == np.uint32(idk)
        Pd = seg(dataset, "points")[ok].astype(np.float64) @ R.T + t
        Pm = seg(model, "points")[ok].astype(np.float64)
        Nm = seg(model, "normals")[ok].astype(np.float64)
        r = np.einsum("ij,ij->i", Pd - Pm, Nm)
        keep = np.abs(r) < max_dist
        Pd, Nm, r = Pd[keep], Nm[keep], r[keep]
        J = np.concatenate([np.cross(Pd, Nm), Nm], axis=1)
        w = np.ones_like(r) if kernel_scale <= 0 else kernel_scale ** 2 / (kernel_scale + r) ** 2
        out.append({"JTwJ": (J * w[:, None]).T @ J, "JTwr": (J * w[:, None]).T @ r,
                    "wr2": float(np.sum(w * r * r)), "n_meas": int(len(r))})
    return out
