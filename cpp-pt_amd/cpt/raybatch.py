"""Seeded ray batches for closest-hit / any-hit parity checks.

Mix: 50% camera rays, 45% cosine-distributed rays leaving random surface points
(offset by EPS along the geometric normal, like the reference's spawn offset,
pathtracerMIS.h:126), 5% adversarial rays (axis-aligned directions, rays grazing
triangle planes, rays aimed at shared edges and vertices).  Rays: n x 7 float32
(ox oy oz dx dy dz tmax).
"""
from __future__ import annotations

import numpy as np


def _normalize(v):
    return v / np.linalg.norm(v, axis=1, keepdims=True)


def make_rays(scene, n, seed=4, tmax=np.inf, shadow=False):
    a = scene.c_arrays()
    P, I = a["positions"].astype(np.float64), a["indices"]
    cam = a["camera"]
    rng = np.random.default_rng(seed)
    n_cam = n // 2
    n_surf = (n * 45) // 100
    n_adv = n - n_cam - n_surf
    rays = np.zeros((n, 7), np.float32)
    # camera rays
    cx = rng.uniform(-1, 1, n_cam)
    cy = rng.uniform(-1, 1, n_cam)
    X, Y, Z = cam[3:6], cam[6:9], cam[9:12]
    d = (cx[:, None] * cam[12] * cam[13]) * X + (cy[:, None] * cam[12]) * Y - Z
    rays[:n_cam, 0:3] = cam[0:3]
    rays[:n_cam, 3:6] = _normalize(d)
    # surface rays
    tri = rng.integers(0, len(I), n_surf)
    u = rng.random(n_surf)
    v = rng.random(n_surf)
    flip = u + v > 1
    u[flip], v[flip] = 1 - u[flip], 1 - v[flip]
    A, B, C = P[I[tri, 0]], P[I[tri, 1]], P[I[tri, 2]]
    pts = A + (B - A) * u[:, None] + (C - A) * v[:, None]
    ng = np.cross(B - A, C - A)
    ng = ng / np.maximum(np.linalg.norm(ng, axis=1, keepdims=True), 1e-30)
    side = np.where(rng.random(n_surf) < 0.5, 1.0, -1.0)[:, None]
    ng = ng * side
    r1, r2 = rng.random(n_surf), rng.random(n_surf)
    phi = 2 * np.pi * r1
    local = np.stack([np.cos(phi) * np.sqrt(r2), np.sin(phi) * np.sqrt(r2), np.sqrt(1 - r2)], 1)
    t1 = _normalize(np.cross(np.where(np.abs(ng[:, :1]) < 0.9, [[1.0, 0, 0]], [[0, 1.0, 0]]), ng))
    t2 = np.cross(ng, t1)
    dirs = local[:, :1] * t1 + local[:, 1:2] * t2 + local[:, 2:3] * ng
    rays[n_cam:n_cam + n_surf, 0:3] = pts + ng * 1e-4
    rays[n_cam:n_cam + n_surf, 3:6] = _normalize(dirs)
    # adversarial rays
    k0 = n_cam + n_surf
    lo, hi = P.min(0), P.max(0)
    origins = lo + (hi - lo) * rng.random((n_adv, 3))
    kind = rng.integers(0, 3, n_adv)
    axes = np.eye(3)[rng.integers(0, 3, n_adv)] * np.where(rng.random((n_adv, 1)) < 0.5, 1.0, -1.0)
    tri = rng.integers(0, len(I), n_adv)
    A, B, C = P[I[tri, 0]], P[I[tri, 1]], P[I[tri, 2]]
    edge_pt = A + (B - A) * rng.random((n_adv, 1))  # point on a (possibly shared) edge
    to_edge = _normalize(edge_pt - origins)
    graze = _normalize((B - A) + 1e-3 * rng.standard_normal((n_adv, 3)))
    dirs = np.where(kind[:, None] == 0, axes, np.where(kind[:, None] == 1, to_edge, graze))
    org = np.where(kind[:, None] == 2, A + (C - A) * 0.25 - graze * 0.5, origins)
    rays[k0:, 0:3] = org
    rays[k0:, 3:6] = _normalize(dirs)
    rays[:, 6] = tmax
    if shadow:
        # finite segments: tmax uniform in (0.1, 3) scene-scale
        rays[:, 6] = rng.uniform(0.05, 1.0, n) * float(np.linalg.norm(hi - lo))
    return np.ascontiguousarray(rays, np.float32)
