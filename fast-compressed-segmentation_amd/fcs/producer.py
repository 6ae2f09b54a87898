"""Synthetic volumes and the CPU .csdb producer (include/fcs/fcs_producer.h)."""
import ctypes
import os

import numpy as np

from ._lib import EncodeOpts, GenParams, c_u8p, c_u32p, c_u64p, c_vp, check, lib


class Generator:
    """Deterministic label volume recipe (configs 1-5)."""

    def __init__(self, config=None, size=0, params=None):
        p = GenParams()
        if params is not None:
            for k, v in params.items():
                setattr(p, k, v)
        else:
            check(lib.fcs_gen_preset(int(config), int(size), ctypes.byref(p)))
        h = c_vp()
        check(lib.fcs_gen_create(ctypes.byref(p), ctypes.byref(h)))
        self.h = h
        self.params = {f: getattr(p, f) for f, _ in GenParams._fields_}
        self.size = p.size

    def __del__(self):
        if getattr(self, "h", None):
            lib.fcs_gen_destroy(self.h)
            self.h = None

    def box(self, x0, y0, z0, nx, ny, nz, threads=0):
        out = np.zeros(nx * ny * nz, np.uint32)
        check(lib.fcs_gen_box(self.h, x0, y0, z0, nx, ny, nz, out.ctypes.data_as(c_u32p), threads))
        return out.reshape(nz, ny, nx)

    def bricks(self, b, first, count, threads=0):
        """Source labels of bricks first..first+count-1 (Morton IDs), each
        brick's b^3 labels in Morton order: shape (count, b^3)."""
        out = np.zeros((int(count), int(b) ** 3), np.uint32)
        check(lib.fcs_gen_bricks(self.h, int(b), int(first), int(count), out.ctypes.data_as(c_u32p),
                                 _threads(threads)))
        return out

    def volume(self, threads=0):
        n = self.size
        return self.box(0, 0, 0, n, n, n, threads)

    def background_fraction(self):
        return lib.fcs_gen_background_fraction(self.h)


def _threads(t):
    """Explicit count, else OMP_NUM_THREADS when set (the share of a shared
    machine), else the CPUs this process may run on."""
    if t > 0:
        return t
    env = os.environ.get("OMP_NUM_THREADS", "")
    if env.isdigit() and int(env) > 0:
        return int(env)
    try:
        return len(os.sched_getaffinity(0))
    except AttributeError:
        return os.cpu_count() or 4


def encode_volume(labels, brick_size, header_brick_size_64=False, threads=0):
    """Encode a (D, H, W) label array (x fastest) to .csdb bytes."""
    a = np.ascontiguousarray(labels)
    if a.dtype not in (np.uint32, np.uint64):
        raise TypeError("labels must be uint32 or uint64")
    D, H, W = a.shape
    o = EncodeOpts(int(brick_size), 1 if header_brick_size_64 else 0, _threads(threads))
    buf = c_u8p()
    n = ctypes.c_size_t()
    check(lib.fcs_encode_volume(a.ctypes.data_as(c_vp), 64 if a.dtype == np.uint64 else 32, W, H, D,
                                ctypes.byref(o), ctypes.byref(buf), ctypes.byref(n)))
    try:
        return ctypes.string_at(buf, n.value)
    finally:
        lib.fcs_free_buffer(buf)


def encode_generated(gen, brick_size, header_brick_size_64=False, threads=0, progress=False):
    """Encode a generated cube brick by brick (the volume is never materialised).
    progress: the encoder reports every sixteenth of the bricks on stderr."""
    o = EncodeOpts(int(brick_size), 1 if header_brick_size_64 else 0, _threads(threads), 1 if progress else 0)
    buf = c_u8p()
    n = ctypes.c_size_t()
    check(lib.fcs_encode_generated(gen.h, ctypes.byref(o), ctypes.byref(buf), ctypes.byref(n)))
    try:
        return ctypes.string_at(buf, n.value)
    finally:
        lib.fcs_free_buffer(buf)


def brick_pyramid(brick_labels_morton, brick_size):
    """Majority-vote LOD pyramid of one brick (coarse-first layout)."""
    a = np.ascontiguousarray(brick_labels_morton, dtype=np.uint64)
    L = int(brick_size).bit_length() - 1
    out = np.zeros(((1 << (3 * (L + 1))) - 1) // 7, np.uint64)
    check(lib.fcs_brick_pyramid(a.ctypes.data_as(c_u64p), int(brick_size), out.ctypes.data_as(c_u64p), None))
    return out
