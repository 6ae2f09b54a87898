"""Serving over include/fcs/fcs_serve.h: views (asynchronous frames with
device-side LOD selection, optionally replayed as hipGraphs), cache admission
under an HBM budget, the open-loop request loop, and the synthetic traffic of
many viewers (Poisson arrivals, random-walk orbit cameras)."""
import ctypes
import math

import numpy as np

from ._lib import (
    VIEW_CHECKSUM,
    VIEW_GRAPH,
    VIEW_LABELS,
    VIEW_RGBA,
    CameraStruct,
    ServeRequestStruct,
    ServeResultStruct,
    ViewResultStruct,
    c_u8p,
    c_u32p,
    c_u64p,
    c_vp,
    check,
    lib,
)
from .api import Frame

RESOLUTIONS = ((640, 480), (1280, 720), (1920, 1080))


class View:
    """A render target with its own HIP stream (fcs_view_create)."""

    def __init__(self, ctx, max_w=1920, max_h=1080, graph=True, labels=False, checksum=True, rgba=False):
        self.ctx = ctx
        self.rgba = rgba
        flags = (VIEW_GRAPH if graph else 0) | (VIEW_LABELS if labels else 0) | (VIEW_CHECKSUM if checksum else 0)
        flags |= VIEW_RGBA if rgba else 0
        h = c_vp()
        check(lib.fcs_view_create(ctx.h, max_w, max_h, flags, ctypes.byref(h)), ctx.h)
        self.h = h
        self._vol = None

    def close(self):
        if self.h:
            lib.fcs_view_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def render_async(self, vol, cam, w, h, brick_lod=None):
        lp = None
        if brick_lod is not None:
            self._lod = np.ascontiguousarray(brick_lod, dtype=np.uint8)
            lp = self._lod.ctypes.data_as(c_u8p)
        check(lib.fcs_view_render_async(self.h, vol.h, ctypes.byref(cam.s), lp, w, h), self.ctx.h)
        self._vol = vol

    def query(self):
        return check(lib.fcs_view_query(self.h), self.ctx.h) == 1

    def wait(self):
        r = ViewResultStruct()
        check(lib.fcs_view_wait(self.h, ctypes.byref(r)), self.ctx.h)
        return {f: getattr(r, f) for f, _ in ViewResultStruct._fields_ if f != "pad"}

    def labels(self):
        """Copy of the pinned image of the last frame (label views only): the
        first-hit labels, or the RGBA8 words of a shaded view."""
        r = self.wait()
        p = lib.fcs_view_labels(self.h)
        if not p:
            return None
        dt = np.uint64 if (r["label_bits"] == 64 and not self.rgba) else np.uint32
        n = r["w"] * r["h"]
        return np.ctypeslib.as_array(ctypes.cast(p, ctypes.POINTER(ctypes.c_uint8)),
                                     shape=(n * dt().itemsize,)).view(dt).copy()

    def read(self):
        """The last frame's device outputs as a Frame (label, voxel, t_hit, brick_lod_used)."""
        r = self.wait()
        n = r["w"] * r["h"]
        dt = np.uint64 if r["label_bits"] == 64 else np.uint32
        label = np.zeros(n, dt)
        voxel = np.zeros(n, np.uint64)
        t_hit = np.zeros(n, np.float32)
        used = np.zeros(self._vol.info["num_bricks"], np.uint8)
        check(lib.fcs_view_read(self.h, label.ctypes.data, voxel.ctypes.data_as(c_u64p),
                                t_hit.ctypes.data_as(ctypes.POINTER(ctypes.c_float)), used.ctypes.data_as(c_u8p)),
              self.ctx.h)
        return Frame(label, voxel, t_hit, None, used, 0, 0.0, r["device_ms"])

    def read_rgba(self):
        """The last frame's RGBA8 image (shaded views)."""
        r = self.wait()
        out = np.zeros(r["w"] * r["h"], np.uint32)
        check(lib.fcs_view_read_rgba(self.h, out.ctypes.data_as(c_u32p)), self.ctx.h)
        return out

    @property
    def graphs(self):
        return int(lib.fcs_view_graphs(self.h))


def plan(full_bytes, demand, budget):
    """fcs_serve_plan: (plan per volume: 2 whole / 1 on-demand pool / 0 none, pool bytes)."""
    fb = np.ascontiguousarray(full_bytes, dtype=np.uint64)
    d = np.ascontiguousarray(demand, dtype=np.float64)
    n = fb.size
    p = np.zeros(n, np.uint32)
    pool = np.zeros(n, np.uint64)
    check(lib.fcs_serve_plan(fb.ctypes.data_as(c_u64p), d.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), n,
                             int(budget), p.ctypes.data_as(c_u32p), pool.ctypes.data_as(c_u64p)))
    return p, pool


def full_pyramid_bytes(vol):
    i = vol.info
    L = i["levels"]
    return i["num_bricks"] * (((1 << (3 * (L + 1))) - 1) // 7) * (i["label_bits"] // 8)


def admit(vols, demand, budget):
    """fcs_serve_admit on the volumes (same context)."""
    n = len(vols)
    hs = (c_vp * n)(*[v.h for v in vols])
    d = np.ascontiguousarray(demand, dtype=np.float64)
    p = np.zeros(n, np.uint32)
    pool = np.zeros(n, np.uint64)
    ctx = vols[0].ctx
    check(lib.fcs_serve_admit(hs, d.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), n, int(budget),
                              p.ctypes.data_as(c_u32p), pool.ctypes.data_as(c_u64p)), ctx.h)
    return p, pool


def run(views, vols, requests):
    """fcs_serve_run over a request array (see make_trace); returns a dict of
    numpy arrays: issue_s, done_s, checksum, device_ms, misses, view."""
    n = len(requests)
    hv = (c_vp * len(views))(*[v.h for v in views])
    hs = (c_vp * len(vols))(*[v.h for v in vols])
    out = (ServeResultStruct * max(n, 1))()
    ctx = vols[0].ctx
    check(lib.fcs_serve_run(hv, len(views), hs, len(vols), requests, n, out), ctx.h)
    dtypes = {"issue_s": np.float64, "done_s": np.float64, "checksum": np.uint64, "device_ms": np.float32,
              "misses": np.uint32, "view": np.uint32}
    return {f: np.fromiter((getattr(out[i], f) for i in range(n)), dtype=dt, count=n) for f, dt in dtypes.items()}


def poisson_arrivals(rate, duration, rng):
    """Arrival times of a Poisson process of `rate` per second over [0, duration)."""
    t = []
    now = rng.exponential(1.0 / rate)
    while now < duration:
        t.append(now)
        now += rng.exponential(1.0 / rate)
    return np.array(t)


def make_trace(vols, n_sessions=64, rate=2000.0, duration=5.0, seed=0, resolutions=RESOLUTIONS,
               radius=(0.8, 1.5), step=(0.03, 0.02, 0.02), fov=math.pi / 3):
    """Open-loop traffic of `n_sessions` viewers: one Poisson stream of total
    `rate` frames/s, each request from a session drawn uniformly; session s
    watches volume s % len(vols); its orbit camera (yaw, pitch, radius as a
    multiple of the volume diagonal) takes one Gaussian random-walk step per
    request (std `step`), pitch kept within +-0.45 pi, radius within `radius`;
    resolution uniform over `resolutions`. Returns (ServeRequestStruct array,
    dict of per-request numpy fields)."""
    rng = np.random.default_rng(seed)
    arr = poisson_arrivals(rate, duration, rng)
    n = len(arr)
    sess = rng.integers(0, n_sessions, n)
    res_i = rng.integers(0, len(resolutions), n)
    state = np.stack([rng.uniform(0, 2 * math.pi, n_sessions), rng.uniform(-0.3 * math.pi, 0.3 * math.pi, n_sessions),
                      rng.uniform(radius[0], radius[1], n_sessions)], 1)
    reqs = (ServeRequestStruct * max(n, 1))()
    vol_of = np.zeros(n, np.int64)
    for i in range(n):
        s = int(sess[i])
        st = state[s]
        st[0] += rng.normal(0, step[0])
        st[1] = min(max(st[1] + rng.normal(0, step[1]), -0.45 * math.pi), 0.45 * math.pi)
        st[2] = min(max(st[2] * math.exp(rng.normal(0, step[2])), radius[0]), radius[1])
        vi = s % len(vols)
        v = vols[vi]
        diag = math.sqrt(3.0) * v.info["grid_size"]
        cam = v.orbit_camera(float(st[0]), float(st[1]), float(st[2]) * diag, fov)
        w, h = resolutions[int(res_i[i])]
        r = reqs[i]
        r.arrival_s = float(arr[i])
        r.session = s
        r.volume = vi
        r.w = w
        r.h = h
        ctypes.memmove(ctypes.addressof(r.cam), ctypes.addressof(cam.s), ctypes.sizeof(CameraStruct))
        vol_of[i] = vi
    meta = {"arrival_s": arr, "session": sess, "volume": vol_of,
            "w": np.array([resolutions[int(k)][0] for k in res_i]), "h": np.array([resolutions[int(k)][1] for k in res_i])}
    return reqs, meta


def request_camera(req):
    """The fcs camera of one ServeRequestStruct (for re-rendering it)."""
    from .api import Camera
    c = Camera((0, 0, 0), (0, 0, 1), (1, 0, 0), (0, 1, 0))
    ctypes.memmove(ctypes.addressof(c.s), ctypes.addressof(req.cam), ctypes.sizeof(CameraStruct))
    return c


def latency_stats(meta, res):
    lat = (res["done_s"] - meta["arrival_s"]) * 1e3
    span = float(res["done_s"].max() - meta["arrival_s"].min()) if len(lat) else 0.0
    px = meta["w"] * meta["h"]
    return {
        "frames": int(len(lat)),
        "p50_ms": float(np.percentile(lat, 50)) if len(lat) else 0.0,
        "p99_ms": float(np.percentile(lat, 99)) if len(lat) else 0.0,
        "mean_ms": float(lat.mean()) if len(lat) else 0.0,
        "max_ms": float(lat.max()) if len(lat) else 0.0,
        "frames_per_s": len(lat) / span if span > 0 else 0.0,
        "rays_per_s": float(px.sum()) / span if span > 0 else 0.0,
        "device_ms_mean": float(res["device_ms"].mean()) if len(lat) else 0.0,
        "missed_frames": int(np.count_nonzero(res["misses"])),
    }
