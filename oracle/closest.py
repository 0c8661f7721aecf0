"""Closest-point oracle (test infrastructure): FP64 brute force over all triangles of the
closest point of a triangle (Ericson, Real-Time Collision Detection, 5.1.5), the query
that EmbreeScene::closestPoint answers (src/rmagine_embree/src/map/embree/EmbreeScene.cpp:292-315,
callback EmbreeMesh.cpp:22-131): nearest triangle strictly within max_distance, equal
distances to the lower face index.  Pinned in tests/test_closest_oracle.py by closed forms
(cube faces, edges, corners), the Voronoi regions of one triangle, and dense sampling.
Only tests import this module."""
from __future__ import annotations

import numpy as np

FLT_MAX = float(np.finfo(np.float32).max)


def closest_on_triangles(p, A, B, C):
    """Closest points of triangles (A, B, C) ((T,3) float64 each) to the point p (3,)."""
    p = np.asarray(p, np.float64)
    ab, ac, ap = B - A, C - A, p - A
    d1, d2 = (ab * ap).sum(1), (ac * ap).sum(1)
    bp = p - B
    d3, d4 = (ab * bp).sum(1), (ac * bp).sum(1)
    cp = p - C
    d5, d6 = (ab * cp).sum(1), (ac * cp).sum(1)
    vc = d1 * d4 - d3 * d2
    vb = d5 * d2 - d1 * d6
    va = d3 * d6 - d5 * d4
    with np.errstate(divide="ignore", invalid="ignore"):
        den = 1.0 / (va + vb + vc)
        out = A + ab * (vb * den)[:, None] + ac * (vc * den)[:, None]                       # face
        bc_t = (d4 - d3) / ((d4 - d3) + (d5 - d6))
        m = (va <= 0) & ((d4 - d3) >= 0) & ((d5 - d6) >= 0)
        out[m] = (B + (C - B) * bc_t[:, None])[m]                                          # edge BC
        m_ac = (vb <= 0) & (d2 >= 0) & (d6 <= 0)
        out[m_ac] = (A + ac * (d2 / (d2 - d6))[:, None])[m_ac]                             # edge AC
        m_c = (d6 >= 0) & (d5 <= d6)
        out[m_c] = C[m_c]                                                                  # vertex C
        m_ab = (vc <= 0) & (d1 >= 0) & (d3 <= 0)
        out[m_ab] = (A + ab * (d1 / (d1 - d3))[:, None])[m_ab]                             # edge AB
        m_b = (d3 >= 0) & (d4 <= d3)
        out[m_b] = B[m_b]                                                                  # vertex B
        m_a = (d1 <= 0) & (d2 <= 0)
        out[m_a] = A[m_a]                                                                  # vertex A
    return out


def closest_points(V, F, Q, max_distance=FLT_MAX):
    """dist (float32), point (float32), face index (int64, -1 for none) of every query."""
    V = np.asarray(V, np.float32).astype(np.float64)
    F = np.asarray(F, np.int64)
    A, B, C = V[F[:, 0]], V[F[:, 1]], V[F[:, 2]]
    Q = np.asarray(Q, np.float32).astype(np.float64)
    best0 = float(np.float32(max_distance)) ** 2
    dist = np.full(len(Q), FLT_MAX, np.float32)
    pts = np.full((len(Q), 3), np.nan, np.float32)
    face = np.full(len(Q), -1, np.int64)
    for i, q in enumerate(Q):
        P = closest_on_triangles(q, A, B, C)
        d2 = ((P - q) ** 2).sum(1)
        j = int(np.argmin(d2))  # first (lowest face index) among equal distances
        if d2[j] < best0:
            dist[i] = np.float32(np.sqrt(d2[j]))
            pts[i] = P[j].astype(np.float32)
            face[i] = j
    return dist, pts, face
