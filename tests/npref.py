"""Independent NumPy checks used to pin the oracle (test infrastructure).

closest_hits_mt: brute-force closest hit with the Moller-Trumbore ray/triangle test in
float64 ("Fast, minimum storage ray-triangle intersection", JGT 1997), a different
algorithm than the oracle's watertight Woop test, vectorised over triangles.
rays_for_model: FP32 ray directions/origins restated from
src/rmagine_core/include/rmagine/types/sensor_models.h:170-175,256-277,376-385 with
NumPy float32 arithmetic (no FMA), and poses composed with the quaternion formulas of
Quaternion.tcc:35-75 / Transform.tcc:34-62.
"""
from __future__ import annotations

import math

import numpy as np

f4 = np.float32


def qmul(a, b):
    ax, ay, az, aw = a
    bx, by, bz, bw = b
    return (f4(f4(f4(aw * bx) + f4(ax * bw)) + f4(ay * bz)) - f4(az * by),
            f4(f4(f4(aw * by) - f4(ax * bz)) + f4(ay * bw)) + f4(az * bx),
            f4(f4(f4(aw * bz) + f4(ax * by)) - f4(ay * bx)) + f4(az * bw),
            f4(f4(f4(aw * bw) - f4(ax * bx)) - f4(ay * by)) - f4(az * bz))


def qrot(q, p):
    P = (f4(p[0]), f4(p[1]), f4(p[2]), f4(0.0))
    qi = (-q[0], -q[1], -q[2], q[3])
    r = qmul(qmul(q, P), qi)
    return np.array(r[:3], dtype=np.float32)


def tmul(A, B):
    qa, ta = tuple(map(f4, A[0:4])), np.asarray(A[4:7], dtype=np.float32)
    qb, tb = tuple(map(f4, B[0:4])), np.asarray(B[4:7], dtype=np.float32)
    t = qrot(qa, tb)
    R = qmul(qa, qb)
    t = (t + ta).astype(np.float32)
    return np.array([*R, *t, 0.0], dtype=np.float32)


def sphere_dirs(model):
    H, W = model.phi.size, model.theta.size
    phi = np.array([model.phi.getValue(i) for i in range(H)], dtype=np.float32)
    th = np.array([model.theta.getValue(j) for j in range(W)], dtype=np.float32)
    cp = np.cos(phi.astype(np.float64)).astype(np.float32)
    sp = np.sin(phi.astype(np.float64)).astype(np.float32)
    ct = np.cos(th.astype(np.float64)).astype(np.float32)
    st = np.sin(th.astype(np.float64)).astype(np.float32)
    d = np.empty((H, W, 3), dtype=np.float32)
    d[..., 0] = cp[:, None] * ct[None, :]
    d[..., 1] = cp[:, None] * st[None, :]
    d[..., 2] = np.broadcast_to(sp[:, None], (H, W))
    return d.reshape(-1, 3)


def pinhole_dirs(model):
    W, H = model.width, model.height
    hid = np.arange(W, dtype=np.float32)[None, :].repeat(H, 0)
    vid = np.arange(H, dtype=np.float32)[:, None].repeat(W, 1)
    pX = (hid - f4(model.c[0])) / f4(model.f[0])
    pY = (vid - f4(model.c[1])) / f4(model.f[1])
    z = np.ones_like(pX)
    n = np.sqrt((pX * pX + pY * pY) + z * z).astype(np.float32)
    ox, oy, oz = pX / n, pY / n, z / n
    return np.stack([oz, -ox, -oy], axis=-1).reshape(-1, 3).astype(np.float32)


def closest_hits_mt(V, F, orig, dirs, tmax):
    """Closest hit per ray (float64 Moller-Trumbore, two-sided, 0 <= t <= tmax).

    Returns (t, face) with face = -1 for a miss; ties go to the lower face index.
    """
    V = np.asarray(V, dtype=np.float64)
    F = np.asarray(F, dtype=np.int64)
    v0, v1, v2 = V[F[:, 0]], V[F[:, 1]], V[F[:, 2]]
    e1, e2 = v1 - v0, v2 - v0
    ts = np.full(len(dirs), np.inf)
    fs = np.full(len(dirs), -1, dtype=np.int64)
    for i, (o, d) in enumerate(zip(np.asarray(orig, np.float64), np.asarray(dirs, np.float64))):
        p = np.cross(d, e2)
        det = np.einsum("ij,ij->i", e1, p)
        with np.errstate(divide="ignore", invalid="ignore"):
            inv = 1.0 / det
            s = o - v0
            u = np.einsum("ij,ij->i", s, p) * inv
            q = np.cross(s, e1)
            v = (q @ d) * inv
            t = np.einsum("ij,ij->i", e2, q) * inv
        ok = (det != 0) & (u >= 0) & (v >= 0) & (u + v <= 1) & (t >= 0) & (t <= tmax)
        if ok.any():
            idx = np.nonzero(ok)[0]
            tt = t[idx]
            j = np.lexsort((idx, tt))[0]
            ts[i] = tt[j]
            fs[i] = idx[j]
    return ts, fs


def qrot_many(q, P):
    """qrot of every row of P (float32 (n, 3)) by one quaternion, the same FP32 operations as
    qrot, vectorised."""
    P = np.asarray(P, np.float32)
    qx, qy, qz, qw = (f4(v) for v in q)
    px, py, pz = P[:, 0], P[:, 1], P[:, 2]
    pw = np.zeros_like(px)
    # q * (p, 0)
    ax = ((qw * px + qx * pw) + qy * pz) - qz * py
    ay = ((qw * py - qx * pz) + qy * pw) + qz * px
    az = ((qw * pz + qx * py) - qy * px) + qz * pw
    aw = ((qw * pw - qx * px) - qy * py) - qz * pz
    # (...) * q^-1
    bx, by, bz, bw = -qx, -qy, -qz, qw
    rx = ((aw * bx + ax * bw) + ay * bz) - az * by
    ry = ((aw * by - ax * bz) + ay * bw) + az * bx
    rz = ((aw * bz + ax * by) - ay * bx) + az * bw
    return np.stack([rx, ry, rz], 1).astype(np.float32)
