"""Synthetic meshes and the benchmark workloads.

genCube / genPlane / genSphere reproduce the vertex and face order of Rmagine's
synthetic meshes (src/rmagine_core/src/util/synthetic.cpp:66-251), because face ids of
the reference's known-answer tests refer to that order.  The other generators build the
procedural workloads (room with boxes, fBm terrain, random boxes and icospheres), all
from numpy's PCG64 generator seeded with 1000 + workload index, so every run and every
machine sees the same meshes and poses.

All meshes are returned as (vertices float32 (V,3), faces uint32 (F,3)).
"""
from __future__ import annotations

import math

import numpy as np

from . import types as T

SEED_BASE = 1000


def genCube():
    """Unit cube, 8 vertices, 12 faces (synthetic.cpp:183-221)."""
    v = np.array([[-0.5, -0.5, -0.5], [-0.5, -0.5, 0.5], [-0.5, 0.5, 0.5], [-0.5, 0.5, -0.5],
                  [0.5, 0.5, -0.5], [0.5, 0.5, 0.5], [0.5, -0.5, 0.5], [0.5, -0.5, -0.5]], dtype=np.float32)
    f = np.array([[0, 1, 2], [2, 3, 0], [4, 5, 6], [6, 7, 4],
                  [0, 7, 1], [1, 7, 6], [1, 6, 5], [5, 2, 1], [2, 5, 4], [3, 2, 4], [3, 4, 7], [3, 7, 0]],
                 dtype=np.uint32)
    return v, f


def genPlane():
    """Unit plane in z=0, 4 vertices, 2 faces (synthetic.cpp:232-251)."""
    v = np.array([[-0.5, 0.5, 0.0], [0.5, 0.5, 0.0], [0.5, -0.5, 0.0], [-0.5, -0.5, 0.0]], dtype=np.float32)
    f = np.array([[1, 0, 3], [3, 2, 1]], dtype=np.uint32)
    return v, f


def genSphere(num_long: int = 50, num_lat: int = 100):
    """UV sphere of radius 0.5 (synthetic.cpp:66-164), FP32 vertex arithmetic."""
    radius = np.float32(0.5)
    long_inc = np.float32(math.pi / float(np.float32(num_long + 1)))
    lat_inc = np.float32((2 * math.pi) / float(np.float32(num_lat)))
    verts = [(0.0, 0.0, float(radius))]
    faces = []
    for i in range(num_lat):
        id_bl, id_tl, id_tr = 0, i, i + 1
        if i == num_lat - 1:
            id_tr -= num_lat
        faces.append((id_bl, id_tl + 1, id_tr + 1))
    for i in range(num_long):
        alpha = np.float32(long_inc * np.float32(i + 1))
        sa, ca = np.float32(math.sin(float(alpha))), np.float32(math.cos(float(alpha)))
        for j in range(num_lat):
            beta = np.float32(lat_inc * np.float32(j + 1))
            sb, cb = np.float32(math.sin(float(beta))), np.float32(math.cos(float(beta)))
            verts.append((float(radius * sa * cb), float(radius * sa * sb), float(radius * ca)))
            if i > 0:
                id_bl = num_lat * (i - 1) + j
                id_br = num_lat * (i - 1) + j + 1
                id_tl = num_lat * i + j
                id_tr = num_lat * i + j + 1
                if j == num_lat - 1:
                    id_br -= num_lat
                    id_tr -= num_lat
                faces.append((id_br + 1, id_bl + 1, id_tl + 1))
                faces.append((id_tl + 1, id_tr + 1, id_br + 1))
    verts.append((float(radius * np.float32(math.sin(math.pi)) * np.float32(math.cos(2 * math.pi))),
                  float(radius * np.float32(math.sin(math.pi)) * np.float32(math.sin(2 * math.pi))),
                  float(radius * np.float32(math.cos(math.pi)))))
    nv = len(verts)
    for i in range(nv - 1 - num_lat, nv - 1):
        id_bl, id_br, id_tl = i, i + 1, nv - 1
        if id_br == id_tl:
            id_br -= num_lat
        faces.append((id_br, id_bl, id_tl))
    return np.asarray(verts, dtype=np.float32), np.asarray(faces, dtype=np.uint32)


def icosphere(subdiv: int = 2):
    """Unit icosphere: 20 * 4**subdiv faces (320 for subdiv 2)."""
    t = (1.0 + math.sqrt(5.0)) / 2.0
    v = [(-1, t, 0), (1, t, 0), (-1, -t, 0), (1, -t, 0), (0, -1, t), (0, 1, t), (0, -1, -t), (0, 1, -t),
         (t, 0, -1), (t, 0, 1), (-t, 0, -1), (-t, 0, 1)]
    v = [np.array(p, dtype=np.float64) / np.linalg.norm(p) for p in v]
    f = [(0, 11, 5), (0, 5, 1), (0, 1, 7), (0, 7, 10), (0, 10, 11), (1, 5, 9), (5, 11, 4), (11, 10, 2), (10, 7, 6),
         (7, 1, 8), (3, 9, 4), (3, 4, 2), (3, 2, 6), (3, 6, 8), (3, 8, 9), (4, 9, 5), (2, 4, 11), (6, 2, 10),
         (8, 6, 7), (9, 8, 1)]
    for _ in range(subdiv):
        cache = {}

        def mid(a, b):
            key = (min(a, b), max(a, b))
            if key not in cache:
                m = v[a] + v[b]
                v.append(m / np.linalg.norm(m))
                cache[key] = len(v) - 1
            return cache[key]

        nf = []
        for a, b, c in f:
            ab, bc, ca = mid(a, b), mid(b, c), mid(c, a)
            nf += [(a, ab, ca), (b, bc, ab), (c, ca, bc), (ab, bc, ca)]
        f = nf
    return np.asarray(v, dtype=np.float32), np.asarray(f, dtype=np.uint32)


def _concat(parts):
    vs, fs, off = [], [], 0
    for v, f in parts:
        vs.append(np.asarray(v, dtype=np.float32))
        fs.append(np.asarray(f, dtype=np.uint32) + np.uint32(off))
        off += len(v)
    return np.concatenate(vs).astype(np.float32), np.concatenate(fs).astype(np.uint32)


def _box(lo, hi):
    v, f = genCube()
    lo = np.asarray(lo, dtype=np.float64)
    hi = np.asarray(hi, dtype=np.float64)
    vv = (v.astype(np.float64) + 0.5) * (hi - lo) + lo
    return vv.astype(np.float32), f


def room_with_boxes(n_boxes: int = 200, size=(20.0, 20.0, 4.0), seed: int = SEED_BASE + 0):
    """Box room (12 tris) plus n_boxes axis-aligned boxes standing on the floor.

    Room spans x, y in [-10, 10], z in [0, 4]; box sides U[0.2, 2] m; 2,412 triangles
    for 200 boxes.  Returns (V, F, boxes) with boxes an (n, 6) array of lo/hi corners.
    """
    rng = np.random.default_rng(seed)
    sx, sy, sz = size
    parts = [_box((-sx / 2, -sy / 2, 0.0), (sx / 2, sy / 2, sz))]
    boxes = []
    for _ in range(n_boxes):
        s = rng.uniform(0.2, 2.0, size=3)
        cx = rng.uniform(-sx / 2 + s[0] / 2, sx / 2 - s[0] / 2)
        cy = rng.uniform(-sy / 2 + s[1] / 2, sy / 2 - s[1] / 2)
        lo = (cx - s[0] / 2, cy - s[1] / 2, 0.0)
        hi = (cx + s[0] / 2, cy + s[1] / 2, min(s[2], sz))
        parts.append(_box(lo, hi))
        boxes.append(lo + hi)
    V, F = _concat(parts)
    return V, F, np.asarray(boxes, dtype=np.float64)


def poses_in_room(n: int, boxes: np.ndarray, size=(20.0, 20.0, 4.0), z: float = 1.0, margin: float = 0.3,
                  seed: int = SEED_BASE + 0) -> np.ndarray:
    """Uniform free-space poses at height z with yaw U[-pi, pi] (outside every box)."""
    rng = np.random.default_rng(seed + 7919)
    out = []
    sx, sy, _ = size
    while len(out) < n:
        x = rng.uniform(-sx / 2 + margin, sx / 2 - margin)
        y = rng.uniform(-sy / 2 + margin, sy / 2 - margin)
        yaw = rng.uniform(-math.pi, math.pi)
        inside = ((boxes[:, 0] - margin <= x) & (x <= boxes[:, 3] + margin) &
                  (boxes[:, 1] - margin <= y) & (y <= boxes[:, 4] + margin) & (boxes[:, 5] + margin >= z))
        if inside.any():
            continue
        out.append(np.concatenate([T.quat_from_euler(0.0, 0.0, yaw), [x, y, z, 0.0]]))
    return np.asarray(out, dtype=np.float32)


def vlp16_workload_model() -> T.SphericalModel:
    """VLP-16: 16 rings -15..+15 deg (2 deg), 900 columns (0.4 deg), range 0.1-100 m."""
    return T.SphericalModel(
        phi=T.DiscreteInterval(T.f32(-15.0 * T.DEG_TO_RAD), T.f32(2.0 * T.DEG_TO_RAD), 16),
        theta=T.DiscreteInterval(T.f32(-math.pi), T.f32(0.4 * T.DEG_TO_RAD), 900),
        range=T.Interval(0.1, 100.0))


def ouster_workload_model() -> T.SphericalModel:
    """Ouster-style spherical: 128 rows over +-22.5 deg, 1024 columns over 360 deg, 0.1-100 m."""
    return T.SphericalModel(
        phi=T.DiscreteInterval(T.f32(-22.5 * T.DEG_TO_RAD), T.f32(45.0 / 127.0 * T.DEG_TO_RAD), 128),
        theta=T.DiscreteInterval(T.f32(-math.pi), T.f32(2.0 * math.pi / 1024.0), 1024),
        range=T.Interval(0.1, 100.0))


def _value_noise(n: int, cells: int, rng) -> np.ndarray:
    """Smooth value noise on an n x n grid from a (cells+1)^2 random lattice."""
    lat = rng.uniform(-1.0, 1.0, size=(cells + 1, cells + 1))
    u = np.linspace(0.0, cells, n, endpoint=False)
    i0 = np.floor(u).astype(np.int64)
    fr = u - i0
    s = fr * fr * (3.0 - 2.0 * fr)
    i1 = np.minimum(i0 + 1, cells)
    a = lat[i0][:, i0] * (1 - s)[None, :] + lat[i0][:, i1] * s[None, :]
    b = lat[i1][:, i0] * (1 - s)[None, :] + lat[i1][:, i1] * s[None, :]
    return a * (1 - s)[:, None] + b * s[:, None]


def terrain_heights(n: int = 1024, octaves: int = 6, amplitude: float = 8.0, seed: int = SEED_BASE + 1):
    """fBm heightfield (value noise, lacunarity 2, gain 0.5) scaled to [0, amplitude] m."""
    rng = np.random.default_rng(seed)
    h = np.zeros((n, n), dtype=np.float64)
    amp = 1.0
    cells = 4
    for _ in range(octaves):
        h += amp * _value_noise(n, cells, rng)
        amp *= 0.5
        cells *= 2
    h = (h - h.min()) / (h.max() - h.min()) * amplitude
    return h


def heightfield_mesh(heights: np.ndarray, spacing: float = 0.1):
    """Regular grid mesh: (n*n) vertices, 2*(n-1)^2 triangles, centred on the origin."""
    n = heights.shape[0]
    xs = (np.arange(n, dtype=np.float64) - (n - 1) / 2.0) * spacing
    X, Y = np.meshgrid(xs, xs, indexing="xy")
    V = np.stack([X.ravel(), Y.ravel(), heights.ravel()], axis=1).astype(np.float32)
    i = np.arange(n - 1)
    a = (i[:, None] * n + i[None, :]).ravel()   # row-major: row r (y), col c (x)
    b = a + 1
    c = a + n
    d = a + n + 1
    F = np.empty((2 * len(a), 3), dtype=np.uint32)
    F[0::2] = np.stack([a, b, d], axis=1)
    F[1::2] = np.stack([a, d, c], axis=1)
    return V, F


def terrain_height_at(heights: np.ndarray, spacing: float, x, y):
    n = heights.shape[0]
    gx = np.asarray(x) / spacing + (n - 1) / 2.0
    gy = np.asarray(y) / spacing + (n - 1) / 2.0
    c0 = np.clip(np.floor(gx).astype(np.int64), 0, n - 2)
    r0 = np.clip(np.floor(gy).astype(np.int64), 0, n - 2)
    fx, fy = gx - c0, gy - r0
    h00, h01 = heights[r0, c0], heights[r0, c0 + 1]
    h10, h11 = heights[r0 + 1, c0], heights[r0 + 1, c0 + 1]
    return (h00 * (1 - fx) + h01 * fx) * (1 - fy) + (h10 * (1 - fx) + h11 * fx) * fy


def terrain_workload(n_grid: int = 1024, n_sensors: int = 1000, seed: int = SEED_BASE + 1):
    """Large terrain: n_grid^2 fBm heightfield at 0.1 m, sensors 1.5 m above ground."""
    spacing = 0.1
    h = terrain_heights(n_grid, seed=seed)
    V, F = heightfield_mesh(h, spacing)
    rng = np.random.default_rng(seed + 7919)
    half = (n_grid - 1) / 2.0 * spacing * 0.8
    x = rng.uniform(-half, half, n_sensors)
    y = rng.uniform(-half, half, n_sensors)
    yaw = rng.uniform(-math.pi, math.pi, n_sensors)
    z = terrain_height_at(h, spacing, x, y) + 1.5
    q = np.stack([T.quat_from_euler(0.0, 0.0, a) for a in yaw])
    Tbm = T.make_transforms(np.stack([x, y, z], axis=1), q)
    return V, F, Tbm


def room_workload(n_sensors: int = 100, n_boxes: int = 200, seed: int = SEED_BASE + 0):
    V, F, boxes = room_with_boxes(n_boxes, seed=seed)
    Tbm = poses_in_room(n_sensors, boxes, seed=seed)
    return V, F, Tbm


def _random_rotation(rng) -> np.ndarray:
    q = rng.normal(size=4)
    q /= np.linalg.norm(q)
    x, y, z, w = q
    return np.array([[1 - 2 * (y * y + z * z), 2 * (x * y - w * z), 2 * (x * z + w * y)],
                     [2 * (x * y + w * z), 1 - 2 * (x * x + z * z), 2 * (y * z - w * x)],
                     [2 * (x * z - w * y), 2 * (y * z + w * x), 1 - 2 * (x * x + y * y)]])


def objects_mesh(n_objects: int = 5000, n_ico: int = 1429, extent: float = 200.0, seed: int = SEED_BASE + 2):
    """Random rotated boxes (12 tris) and subdivision-2 icospheres (320 tris) in a cube.

    The defaults give 3,571 boxes + 1,429 icospheres = 500,132 triangles in a 200 m cube,
    object sizes U[0.5, 5] m.
    """
    rng = np.random.default_rng(seed)
    iv, if_ = icosphere(2)
    cv, cf = genCube()
    kinds = np.zeros(n_objects, dtype=bool)
    kinds[rng.choice(n_objects, n_ico, replace=False)] = True
    parts = []
    for is_ico in kinds:
        R = _random_rotation(rng)
        s = rng.uniform(0.5, 5.0, size=3)
        c = rng.uniform(-extent / 2, extent / 2, size=3)
        v, f = (iv, if_) if is_ico else (cv, cf)
        vv = (v.astype(np.float64) * (s / (2.0 if is_ico else 1.0))) @ R.T + c
        parts.append((vv.astype(np.float32), f))
    return _concat(parts)


def random_unit_dirs(n: int, rng) -> np.ndarray:
    d = rng.normal(size=(n, 3))
    return (d / np.linalg.norm(d, axis=1, keepdims=True)).astype(np.float32)


def pinhole_640x480_fov90() -> T.PinholeModel:
    """640 x 480 depth camera, 90 deg horizontal field of view: fx = fy = 320, c = (319.5, 239.5)."""
    return T.PinholeModel(width=640, height=480, range=T.Interval(0.1, 100.0), f=(320.0, 320.0), c=(319.5, 239.5))


def mixed_sensor_workload(n_sensors: int = 2000, extent: float = 200.0, seed: int = SEED_BASE + 2,
                          o1dn_range=(1000, 50000)):
    """Heterogeneous batch: 40 % VLP-16, 30 % pinhole 640x480 (fov 90), 30 % O1Dn with
    N ~ U[1k, 50k] random unit directions; one random pose per sensor inside the cube.

    Returns (V, F, entries) with entries = list of (model, pose (1, 8) float32).
    """
    V, F = objects_mesh(extent=extent, seed=seed)
    rng = np.random.default_rng(seed + 7919)
    n_vlp = int(round(0.4 * n_sensors))
    n_pin = int(round(0.3 * n_sensors))
    n_o1 = n_sensors - n_vlp - n_pin
    vlp = vlp16_workload_model()
    pin = pinhole_640x480_fov90()
    models = [vlp] * n_vlp + [pin] * n_pin
    for _ in range(n_o1):
        n = int(rng.integers(o1dn_range[0], o1dn_range[1] + 1))
        models.append(T.O1DnModel(n, 1, T.Interval(0.1, 100.0), (0.0, 0.0, 0.0), random_unit_dirs(n, rng)))
    entries = []
    for m in models:
        q = rng.normal(size=4)
        q /= np.linalg.norm(q)
        t = rng.uniform(-extent / 2, extent / 2, size=3)
        entries.append((m, T.make_transforms(t, q)))
    return V, F, entries


def tessellated_box(lo, hi, n: int):
    """Axis-aligned box with every face split into an n x n grid (12 n^2 triangles)."""
    lo = np.asarray(lo, np.float64)
    hi = np.asarray(hi, np.float64)
    g = np.linspace(0.0, 1.0, n + 1)
    U, W = np.meshgrid(g, g, indexing="ij")
    parts = []
    for axis in range(3):
        a1, a2 = (axis + 1) % 3, (axis + 2) % 3
        for side in (0.0, 1.0):
            P = np.zeros((n + 1, n + 1, 3))
            P[..., axis] = lo[axis] + side * (hi[axis] - lo[axis])
            P[..., a1] = lo[a1] + U * (hi[a1] - lo[a1])
            P[..., a2] = lo[a2] + W * (hi[a2] - lo[a2])
            idx = np.arange((n + 1) ** 2).reshape(n + 1, n + 1)
            a, b, c, d = idx[:-1, :-1].ravel(), idx[1:, :-1].ravel(), idx[:-1, 1:].ravel(), idx[1:, 1:].ravel()
            f = np.concatenate([np.stack([a, b, d], 1), np.stack([a, d, c], 1)])
            parts.append((P.reshape(-1, 3).astype(np.float32), f.astype(np.uint32)))
    return _concat(parts)


def fine_room_workload(n_particles: int = 10000, seed: int = SEED_BASE + 3):
    """Room with boxes tessellated to ~50k triangles, a true VLP-16 pose and particles
    perturbed by sigma = 0.2 m (x, y) and 3 deg (yaw).

    Room walls 58 x 58 grids (40,368 tris) + 200 boxes 2 x 2 grids (9,600 tris) = 49,968.
    Returns (V, F, T_true (1,8), particles (n,8)).
    """
    rng = np.random.default_rng(seed)
    _, _, boxes = room_with_boxes(seed=SEED_BASE + 0)
    parts = [tessellated_box((-10.0, -10.0, 0.0), (10.0, 10.0, 4.0), 58)]
    for b in boxes:
        parts.append(tessellated_box(b[0:3], b[3:6], 2))
    V, F = _concat(parts)
    T_true = poses_in_room(1, boxes, seed=seed)
    xy = rng.normal(scale=0.2, size=(n_particles, 2))
    yaw = rng.normal(scale=np.deg2rad(3.0), size=n_particles)
    P = np.repeat(T_true, n_particles, 0).copy()
    P[:, 4:6] += xy.astype(np.float32)
    q0 = T_true[0, 0:4].astype(np.float64)
    for i in range(n_particles):
        h = 0.5 * yaw[i]
        qz = np.array([0.0, 0.0, np.sin(h), np.cos(h)])
        x1, y1, z1, w1 = qz
        x2, y2, z2, w2 = q0
        P[i, 0:4] = [w1 * x2 + x1 * w2 + y1 * z2 - z1 * y2, w1 * y2 - x1 * z2 + y1 * w2 + z1 * x2,
                     w1 * z2 + x1 * y2 - y1 * x2 + z1 * w2, w1 * w2 - x1 * x2 - y1 * y2 - z1 * z2]
    return V, F, T_true, P
