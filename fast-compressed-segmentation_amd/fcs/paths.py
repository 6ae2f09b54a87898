"""Camera fly-through paths: Catmull-Rom splines through seeded control points
inside the volume, with the camera basis of camera.js:61-77 (world up +y,
right = forward x up, up = right x forward, computed in float64 and rounded
to float32 like the reference's uniform upload)."""
import math

import numpy as np

from .api import Camera

_MASK = (1 << 64) - 1


def _splitmix(state):
    state = (state + 0x9E3779B97F4A7C15) & _MASK
    z = state
    z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & _MASK
    z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & _MASK
    return state, z ^ (z >> 31)


def control_points(seed, path, n_points, grid, margin=0.1):
    """n_points uniform points in [margin, 1-margin]^3 * grid, keyed by (seed, path)."""
    st = ((seed & 0xFFFFFFFF) << 32 | (path & 0xFFFFFFFF)) ^ 0x5EED0F1CA7
    pts = np.zeros((n_points, 3), np.float64)
    for i in range(n_points):
        for k in range(3):
            st, r = _splitmix(st)
            u = (r >> 11) / float(1 << 53)
            pts[i, k] = (margin + (1.0 - 2.0 * margin) * u) * grid
    return pts


def catmull_rom(P, frames):
    """Positions and tangents at `frames` parameter values spread uniformly
    over the n-1 segments; the end points are repeated as outer controls."""
    P = np.asarray(P, np.float64)
    n = len(P)
    ext = np.concatenate([P[:1], P, P[-1:]])
    pos = np.zeros((frames, 3))
    tan = np.zeros((frames, 3))
    for f in range(frames):
        u = f / max(frames - 1, 1) * (n - 1)
        i = min(int(u), n - 2)
        t = u - i
        p0, p1, p2, p3 = ext[i], ext[i + 1], ext[i + 2], ext[i + 3]
        a = 2 * p1
        b = p2 - p0
        c = 2 * p0 - 5 * p1 + 4 * p2 - p3
        d = -p0 + 3 * p1 - 3 * p2 + p3
        pos[f] = 0.5 * (a + b * t + c * t * t + d * t * t * t)
        tan[f] = 0.5 * (b + 2 * c * t + 3 * d * t * t)
    return pos, tan


def camera_from(position, forward, fov=math.pi / 3):
    f = np.asarray(forward, np.float64)
    f = f / max(np.linalg.norm(f), 1e-12)
    up = np.array([0.0, 1.0, 0.0]) if abs(f[1]) <= 0.999 else np.array([1.0, 0.0, 0.0])
    r = np.cross(f, up)
    r /= np.linalg.norm(r)
    u = np.cross(r, f)
    u /= np.linalg.norm(u)
    return Camera(position, f, r, u, math.tan(fov / 2))


def flythrough(seed, path, grid, frames=600, n_points=8, fov=math.pi / 3):
    """One path of `frames` cameras looking along the spline's tangent."""
    pos, tan = catmull_rom(control_points(seed, path, n_points, grid), frames)
    return [camera_from(pos[i], tan[i], fov) for i in range(frames)]


def extrapolate(prev, cam, k):
    """The pose k frames ahead at the current velocity (orientation kept):
    only past and present poses are used."""
    p0 = np.array(prev.s.position[:], np.float64)
    p1 = np.array(cam.s.position[:], np.float64)
    d = cam.as_dict()
    out = Camera(p1 + k * (p1 - p0), d["forward"], d["right"], d["up"], d["tan_half_fov"], d["ortho_half_height"])
    return out
