"""Closest-point oracle pins: closed forms on a cube, the seven Voronoi regions of one
triangle, and dense sampling of triangles (no sample may be nearer than the answer)."""
import numpy as np

from oracle import closest as C
from rmagine_amd import scenes


def test_cube_closed_forms():
    V, F = scenes.genCube()  # unit cube centred at the origin
    V = np.asarray(V, np.float32) + np.float32([0, 0, 5])  # the reference's test places it at z = 5
    Q = np.array([[0, 0, 0], [0.1, 0.2, 20.0], [3, 0.1, 5.2], [2, 2, 7], [0.1, -0.2, 5.1], [50, 50, 0]], np.float32)
    want_p = np.array([[0, 0, 4.5], [0.1, 0.2, 5.5], [0.5, 0.1, 5.2], [0.5, 0.5, 5.5], [0.1, -0.5, 5.1], [0.5, 0.5, 4.5]])
    want_d = [4.5, 14.5, 2.5, np.sqrt(1.5 ** 2 * 3), 0.3, np.sqrt(49.5 ** 2 * 2 + 4.5 ** 2)]  # inside: nearest face y = -0.5
    d, p, f = C.closest_points(V, F, Q)
    np.testing.assert_allclose(p, want_p, atol=1e-6)
    np.testing.assert_allclose(d, want_d, rtol=1e-6)
    d, p, f = C.closest_points(V, F, Q, max_distance=4.5)  # strictly within: 4.5 is not
    assert f[0] == -1 and d[0] == C.FLT_MAX and f[2] >= 0 and f[4] >= 0


def test_voronoi_regions_of_one_triangle():
    A, B, Cc = np.array([[0.0, 0, 0]]), np.array([[2.0, 0, 0]]), np.array([[0.0, 2, 0]])
    cases = {  # query -> closest point
        (-1, -1, 1): (0, 0, 0), (3, -1, 0): (2, 0, 0), (-1, 3, 2): (0, 2, 0),
        (1, -2, 0): (1, 0, 0), (-2, 1, 0): (0, 1, 0), (2, 2, 1): (1, 1, 0), (0.5, 0.5, 3): (0.5, 0.5, 0),
    }
    for q, want in cases.items():
        got = C.closest_on_triangles(np.array(q, float), A, B, Cc)[0]
        np.testing.assert_allclose(got, want, atol=1e-12)


def test_against_dense_sampling():
    rng = np.random.default_rng(4)
    for _ in range(20):
        T = rng.normal(size=(3, 3))
        q = rng.normal(size=3) * 2
        p = C.closest_on_triangles(q, T[0:1], T[1:2], T[2:3])[0]
        u, v = np.meshgrid(np.linspace(0, 1, 201), np.linspace(0, 1, 201))
        m = u + v <= 1
        S = T[0] + np.outer(u[m], T[1] - T[0]) + np.outer(v[m], T[2] - T[0])
        ds = np.linalg.norm(S - q, axis=1).min()
        d = np.linalg.norm(p - q)
        assert d <= ds + 1e-12 and ds - d < 0.02 * np.linalg.norm(T[1] - T[0]) + 0.02 * np.linalg.norm(T[2] - T[0])
