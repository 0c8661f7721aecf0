"""Host-side logic that needs no GPU: workload generators, model tables, SO(3) helpers."""
import math

import numpy as np
import torch

from rmagine_amd import scenes
from rmagine_amd import types as T


def test_room_workload_sizes_and_determinism():
    V, F, Tbm = scenes.room_workload(n_sensors=100)
    assert len(F) == 12 + 200 * 12 == 2412
    V2, F2, Tbm2 = scenes.room_workload(n_sensors=100)
    assert np.array_equal(V, V2) and np.array_equal(Tbm, Tbm2)
    assert np.all(np.abs(Tbm[:, 6] - 1.0) < 1e-6)
    np.testing.assert_allclose(np.linalg.norm(Tbm[:, :4], axis=1), 1.0, atol=1e-6)


def test_terrain_mesh_counts():
    V, F, Tbm = scenes.terrain_workload(n_grid=64, n_sensors=5)
    assert len(V) == 64 * 64 and len(F) == 2 * 63 * 63
    assert F.max() < len(V)
    n = 1024
    assert 2 * (n - 1) ** 2 == 2_093_058  # the full-size terrain's triangle count


def test_icosphere_and_cube():
    v, f = scenes.icosphere(2)
    assert len(f) == 320
    np.testing.assert_allclose(np.linalg.norm(v, axis=1), 1.0, atol=1e-6)
    v, f = scenes.genCube()
    assert len(v) == 8 and len(f) == 12


def test_example_models_match_reference_constants():
    m = T.vlp16_900()
    assert m.size() == 14400 and m.range.min == 0.5 and m.range.max == 130.0
    p = T.example_pinhole()
    assert (p.width, p.height, p.c, p.f) == (200, 150, (100.0, 75.0), (100.0, 100.0))
    o = T.example_o1dn()
    np.testing.assert_allclose(np.linalg.norm(o.dirs, axis=1), 1.0, atol=1e-6)


def test_so3_left_jacobian_against_numeric():
    from rmagine_amd.adjoint import perturb, so3_left_jacobian
    rng = np.random.default_rng(0)
    w = torch.as_tensor(rng.normal(scale=0.5, size=(6, 3)), dtype=torch.float64)
    J = so3_left_jacobian(w).numpy()

    def rotvec_quat(v):
        a = np.linalg.norm(v)
        return np.r_[np.sin(a / 2) * v / a, np.cos(a / 2)] if a > 0 else np.r_[0, 0, 0, 1.0]

    def qmul(a, b):
        x1, y1, z1, w1 = a
        x2, y2, z2, w2 = b
        return np.array([w1 * x2 + x1 * w2 + y1 * z2 - z1 * y2, w1 * y2 - x1 * z2 + y1 * w2 + z1 * x2,
                         w1 * z2 + x1 * y2 - y1 * x2 + z1 * w2, w1 * w2 - x1 * x2 - y1 * y2 - z1 * z2])

    h = 1e-6
    for i in range(6):
        for k in range(3):
            e = np.zeros(3)
            e[k] = h
            qa, qb = rotvec_quat(w[i].numpy() + e), rotvec_quat(w[i].numpy() - e)
            rel = qmul(qa, qb * np.array([-1, -1, -1, 1]))  # Exp(w+e) Exp(w-e)^-1 ~ Exp(J (2e))
            rel *= np.sign(rel[3])
            col = 2 * rel[:3] / (2 * h)
            np.testing.assert_allclose(col, J[i][:, k], atol=1e-6)
    # perturb() applies Exp(omega) on the left of R0 and adds tau
    T0 = torch.as_tensor(T.make_transforms([[1, 2, 3]], [T.quat_from_euler(0.1, 0.2, 0.3)]))
    xi = torch.tensor([[0.5, 0.0, -0.5, 0.0, 0.0, math.pi / 2]], dtype=torch.float64)
    Tp = perturb(T0, xi).numpy()[0]
    np.testing.assert_allclose(Tp[4:7], [1.5, 2.0, 2.5], atol=1e-6)
    qz = np.array([0, 0, math.sin(math.pi / 4), math.cos(math.pi / 4)])
    np.testing.assert_allclose(Tp[:4], qmul(qz, T0.numpy()[0, :4].astype(np.float64)), atol=1e-6)
