"""ctypes wrapper of the CPU oracle (oracle.c) and, where built, of the
reference encoder/decoder (oracle/_ref/libref.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, as the checker. The product never imports it.
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ORACLE_SO = os.path.join(HERE, "liboracle.so")
REF_SO = os.path.join(HERE, "_ref", "libref.so")
DDACHECK_SO = os.path.join(HERE, "libddacheck.so")

c_vp = ctypes.c_void_p
c_u8p = ctypes.POINTER(ctypes.c_uint8)
c_u32p = ctypes.POINTER(ctypes.c_uint32)
c_u64p = ctypes.POINTER(ctypes.c_uint64)


class OrBrick(ctypes.Structure):
    _fields_ = [("id", ctypes.c_uint32), ("P", ctypes.c_uint32), ("E", ctypes.c_uint32), ("nsym", ctypes.c_uint32),
                ("pal_off", ctypes.c_uint64), ("str_off", ctypes.c_uint64)]


class OrDataset(ctypes.Structure):
    _fields_ = [("num_bricks", ctypes.c_uint32), ("header_brick_size", ctypes.c_uint32),
                ("brick_size", ctypes.c_uint32), ("version", ctypes.c_uint32),
                ("freq", ctypes.c_uint32 * 16), ("cum", ctypes.c_uint32 * 17), ("total", ctypes.c_uint32),
                ("bricks", ctypes.POINTER(OrBrick)), ("palette", c_u64p), ("src", c_vp)]


class OrCamera(ctypes.Structure):
    _fields_ = [("position", ctypes.c_float * 3), ("tan_half_fov", ctypes.c_float), ("forward", ctypes.c_float * 3),
                ("ortho_half_height", ctypes.c_float), ("right", ctypes.c_float * 3), ("pad1", ctypes.c_float),
                ("up", ctypes.c_float * 3), ("pad2", ctypes.c_float)]


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(ORACLE_SO):
            raise ImportError(f"{ORACLE_SO} missing: run `make -C oracle`")
        L = ctypes.CDLL(ORACLE_SO)
        L.or_parse_csdb.argtypes = [c_vp, ctypes.c_size_t, ctypes.c_uint32, ctypes.POINTER(OrDataset)]
        L.or_free.argtypes = [ctypes.POINTER(OrDataset)]
        L.or_rans_decode.argtypes = [ctypes.POINTER(OrDataset), ctypes.c_uint32, c_u8p, c_u32p]
        L.or_rans_decode.restype = ctypes.c_size_t
        L.or_decode_brick.argtypes = [c_u8p, ctypes.c_size_t, c_u64p, ctypes.c_uint32, ctypes.c_uint32, c_u64p,
                                      c_u32p]
        L.or_decode_dataset_brick.argtypes = [ctypes.POINTER(OrDataset), ctypes.c_uint32, ctypes.c_uint32, c_u64p,
                                              c_u32p]
        L.or_render.argtypes = [ctypes.c_uint32, ctypes.c_uint32, c_u8p, c_u8p, ctypes.POINTER(c_u64p), c_u64p,
                                ctypes.POINTER(OrCamera), ctypes.c_uint32, ctypes.c_uint32, c_u64p, c_u64p,
                                ctypes.POINTER(ctypes.c_float), c_u32p]
        L.or_render_rgba.argtypes = [ctypes.c_uint32, ctypes.c_uint32, c_u8p, c_u8p, ctypes.POINTER(c_u64p), c_u64p,
                                     ctypes.POINTER(OrCamera), ctypes.c_uint32, ctypes.c_uint32, c_u64p, c_u32p,
                                     ctypes.c_uint64, ctypes.POINTER(ctypes.c_float), ctypes.c_float, ctypes.c_float,
                                     c_u32p]
        L.or_brick_lod.argtypes = [ctypes.c_double] * 9 + [ctypes.c_uint32]
        L.or_brick_lod.restype = ctypes.c_uint32
        L.or_decode_all_pyramid.argtypes = [ctypes.POINTER(OrDataset), c_u64p, ctypes.c_int, c_u32p]
        L.or_decode_all_pyramid.restype = ctypes.c_uint64
        L.or_decode_requests.argtypes = [ctypes.POINTER(OrDataset), c_u32p, c_u8p, ctypes.c_uint32, c_u64p, c_u32p,
                                         c_u64p, ctypes.c_int, c_u32p]
        _lib = L
    return _lib


def pyramid_offset(t):
    return ((1 << (3 * t)) - 1) // 7


class Dataset:
    """Parsed .csdb (or_parse_csdb); bricks sorted by Morton ID."""

    def __init__(self, data, brick_size_override=0):
        self.buf = np.frombuffer(bytes(data), dtype=np.uint8).copy()
        self.d = OrDataset()
        rc = lib().or_parse_csdb(self.buf.ctypes.data_as(c_vp), self.buf.size, brick_size_override,
                                 ctypes.byref(self.d))
        if rc != 0:
            raise ValueError("Invalid file format")
        self.num_bricks = self.d.num_bricks
        self.brick_size = self.d.brick_size
        self.L = self.brick_size.bit_length() - 1

    def __del__(self):
        try:
            lib().or_free(ctypes.byref(self.d))
        except Exception:
            pass

    def brick(self, i):
        return self.d.bricks[i]

    def palette(self, i):
        b = self.d.bricks[i]
        return np.ctypeslib.as_array(self.d.palette, shape=(b.pal_off + b.P,))[b.pal_off:].copy()

    def rans(self, i):
        b = self.d.bricks[i]
        out = np.zeros((b.nsym + 2) // 2 + 1, np.uint8)
        st = ctypes.c_uint32()
        n = lib().or_rans_decode(ctypes.byref(self.d), i, out.ctypes.data_as(c_u8p), ctypes.byref(st))
        return out[:n], st.value

    def decode(self, i, t):
        out = np.zeros(1 << (3 * t), np.uint64)
        st = ctypes.c_uint32()
        lib().or_decode_dataset_brick(ctypes.byref(self.d), i, t, out.ctypes.data_as(c_u64p), ctypes.byref(st))
        return out, st.value

    def decode_all_pyramid(self, threads=0):
        per = pyramid_offset(self.L + 1)
        out = np.zeros(per * self.num_bricks, np.uint64)
        st = ctypes.c_uint32()
        lib().or_decode_all_pyramid(ctypes.byref(self.d), out.ctypes.data_as(c_u64p), threads or (os.cpu_count() or 1),
                                    ctypes.byref(st))
        return out.reshape(self.num_bricks, per), st.value


    def decode_requests(self, bricks, lods, label_dtype=np.uint64, threads=0):
        """or_decode_requests: (labels of every request concatenated in
        request order, start offset of each request, status words)."""
        b = np.ascontiguousarray(bricks, dtype=np.uint32)
        t = np.ascontiguousarray(lods, dtype=np.uint8)
        sizes = np.left_shift(np.uint64(1), np.uint64(3) * t.astype(np.uint64))
        off = np.zeros(b.size, np.uint64)
        if b.size:
            off[1:] = np.cumsum(sizes)[:-1]
        total = int(sizes.sum()) if b.size else 0
        out = np.zeros(total, label_dtype)
        st = np.zeros(b.size, np.uint32)
        wide = np.dtype(label_dtype) == np.uint64
        lib().or_decode_requests(ctypes.byref(self.d), b.ctypes.data_as(c_u32p), t.ctypes.data_as(c_u8p), b.size,
                                 off.ctypes.data_as(c_u64p), None if wide else out.ctypes.data_as(c_u32p),
                                 out.ctypes.data_as(c_u64p) if wide else None, threads or (os.cpu_count() or 1),
                                 st.ctypes.data_as(c_u32p))
        return out, off, st


def decode_nibbles(nib_bytes, palette, t):
    out = np.zeros(1 << (3 * t), np.uint64)
    st = ctypes.c_uint32()
    nb = np.ascontiguousarray(nib_bytes, dtype=np.uint8)
    pal = np.ascontiguousarray(palette, dtype=np.uint64)
    lib().or_decode_brick(nb.ctypes.data_as(c_u8p), nb.size, pal.ctypes.data_as(c_u64p), pal.size, t,
                          out.ctypes.data_as(c_u64p), ctypes.byref(st))
    return out, st.value


def render(bricks_per_axis, brick_size, brick_empty, brick_lod, brick_labels, palette0, cam, w, h):
    """or_render. brick_labels: list of uint64 arrays (LOD-t labels, Morton order)
    or None per brick; cam: dict with position/forward/right/up/tan_half_fov/ortho_half_height."""
    nb = len(brick_lod)
    ptrs = (c_u64p * nb)()
    keep = []
    for i, a in enumerate(brick_labels):
        if a is None:
            ptrs[i] = c_u64p()
        else:
            a = np.ascontiguousarray(a, dtype=np.uint64)
            keep.append(a)
            ptrs[i] = a.ctypes.data_as(c_u64p)
    c = OrCamera()
    c.position[:] = cam["position"]
    c.forward[:] = cam["forward"]
    c.right[:] = cam["right"]
    c.up[:] = cam["up"]
    c.tan_half_fov = cam["tan_half_fov"]
    c.ortho_half_height = cam["ortho_half_height"]
    empty = np.ascontiguousarray(brick_empty, dtype=np.uint8)
    lods = np.ascontiguousarray(brick_lod, dtype=np.uint8)
    p0 = np.ascontiguousarray(palette0, dtype=np.uint64)
    lab = np.zeros(w * h, np.uint64)
    vox = np.zeros(w * h, np.uint64)
    tt = np.zeros(w * h, np.float32)
    vis = np.zeros(nb, np.uint32)
    lib().or_render(bricks_per_axis, brick_size, empty.ctypes.data_as(c_u8p), lods.ctypes.data_as(c_u8p), ptrs,
                    p0.ctypes.data_as(c_u64p), ctypes.byref(c), w, h, lab.ctypes.data_as(c_u64p),
                    vox.ctypes.data_as(c_u64p), tt.ctypes.data_as(ctypes.POINTER(ctypes.c_float)),
                    vis.ctypes.data_as(c_u32p))
    return lab, vox, tt, vis


def _ptrs(brick_labels):
    ptrs = (c_u64p * len(brick_labels))()
    keep = []
    for i, a in enumerate(brick_labels):
        if a is None:
            ptrs[i] = c_u64p()
        else:
            a = np.ascontiguousarray(a, dtype=np.uint64)
            keep.append(a)
            ptrs[i] = a.ctypes.data_as(c_u64p)
    return ptrs, keep


def _camera(cam):
    c = OrCamera()
    c.position[:] = cam["position"]
    c.forward[:] = cam["forward"]
    c.right[:] = cam["right"]
    c.up[:] = cam["up"]
    c.tan_half_fov = cam["tan_half_fov"]
    c.ortho_half_height = cam["ortho_half_height"]
    return c


def render_rgba(bricks_per_axis, brick_size, brick_empty, brick_lod, brick_labels, palette0, cam, w, h, tf_labels,
                tf_rgba, light, ambient=0.5, diffuse=0.8):
    """or_render_rgba: the shaded composite, w*h RGBA8 (uint32)."""
    ptrs, keep = _ptrs(brick_labels)
    c = _camera(cam)
    empty = np.ascontiguousarray(brick_empty, dtype=np.uint8)
    lods = np.ascontiguousarray(brick_lod, dtype=np.uint8)
    p0 = np.ascontiguousarray(palette0, dtype=np.uint64)
    tl = np.ascontiguousarray(tf_labels, dtype=np.uint64)
    tr = np.ascontiguousarray(tf_rgba, dtype=np.uint32)
    li = (ctypes.c_float * 3)(*[float(x) for x in light])
    out = np.zeros(w * h, np.uint32)
    lib().or_render_rgba(bricks_per_axis, brick_size, empty.ctypes.data_as(c_u8p), lods.ctypes.data_as(c_u8p),
                         ptrs, p0.ctypes.data_as(c_u64p), ctypes.byref(c), w, h, tl.ctypes.data_as(c_u64p),
                         tr.ctypes.data_as(c_u32p), tl.size, li, ctypes.c_float(ambient), ctypes.c_float(diffuse),
                         out.ctypes.data_as(c_u32p))
    return out


def brick_lod(center, brick_size, cam_pos, fov, screen_width, max_lod):
    return lib().or_brick_lod(*center, brick_size, *cam_pos, fov, screen_width, max_lod)


# ------------------------------------------------------------ reference --
_ref = None


def ref_available():
    return os.path.exists(REF_SO)


def ref():
    global _ref
    if _ref is None:
        R = ctypes.CDLL(REF_SO)
        R.ref_encode.argtypes = [c_u64p, ctypes.c_uint32, ctypes.c_uint32, ctypes.c_uint32, ctypes.c_uint32,
                                 ctypes.c_char_p, c_u64p]
        R.ref_load.argtypes = [ctypes.c_char_p]
        R.ref_rans.argtypes = [ctypes.c_uint32, c_u8p]
        R.ref_brick_id.argtypes = [ctypes.c_uint32]
        R.ref_brick_id.restype = ctypes.c_uint32
        R.ref_decode_finest.argtypes = [ctypes.c_uint32, ctypes.c_uint32, c_u64p]
        R.ref_decode_finest_batch.argtypes = [c_u32p, ctypes.c_uint32, ctypes.c_uint32, ctypes.c_int, c_u64p, c_u32p]
        R.ref_decode_finest_batch.restype = ctypes.c_int64
        _ref = R
    return _ref


def ref_encode(vol_zyx, brick_size, path, with_pyramid=False):
    """Reference encoder on a (D,H,W) volume; returns (.csbd bytes, pyramids in
    file order or None)."""
    v = np.ascontiguousarray(vol_zyx, dtype=np.uint64)
    D, H, W = v.shape
    nbr = -(-W // brick_size) * -(-H // brick_size) * -(-D // brick_size)
    L = brick_size.bit_length() - 1
    pyr = np.zeros(nbr * pyramid_offset(L + 1), np.uint64) if with_pyramid else None
    n = ref().ref_encode(v.ctypes.data_as(c_u64p), W, H, D, brick_size, path.encode(),
                         pyr.ctypes.data_as(c_u64p) if pyr is not None else None)
    if n < 0:
        raise RuntimeError(f"reference encoder failed ({n})")
    with open(path, "rb") as f:
        data = f.read()
    return data, (pyr.reshape(n, -1) if pyr is not None else None)


# ------------------------------------------------ float64 DDA check --
class DcStats(ctypes.Structure):
    _fields_ = [("pixels", ctypes.c_uint64), ("hits", ctypes.c_uint64), ("agree", ctypes.c_uint64),
                ("grazing", ctypes.c_uint64), ("unexplained", ctypes.c_uint64), ("max_t_err", ctypes.c_double)]


_dc = None


def dc():
    global _dc
    if _dc is None:
        if not os.path.exists(DDACHECK_SO):
            raise ImportError(f"{DDACHECK_SO} missing: run `make -C oracle`")
        D = ctypes.CDLL(DDACHECK_SO)
        D.dc_check_frame.argtypes = [ctypes.c_uint32, ctypes.c_uint32, c_u8p, c_u8p, ctypes.POINTER(c_u64p), c_u64p,
                                     ctypes.POINTER(OrCamera), ctypes.c_uint32, ctypes.c_uint32, c_u64p, c_u64p,
                                     ctypes.POINTER(ctypes.c_float), ctypes.c_int, c_u64p, c_u64p, c_u8p,
                                     ctypes.POINTER(DcStats)]
        _dc = D
    return _dc


def check_frame(bricks_per_axis, brick_size, brick_empty, brick_lod, brick_labels, palette0, cam, w, h, label, voxel,
                t_hit, threads=0):
    """dc_check_frame: the float64 exact traversal against a first-hit frame.
    Returns (stats dict, per-pixel class array, exact labels, exact voxels);
    classes: 0 agree, 1 grazing tie, 2 unexplained."""
    ptrs, keep = _ptrs(brick_labels)
    c = _camera(cam)
    empty = np.ascontiguousarray(brick_empty, dtype=np.uint8)
    lods = np.ascontiguousarray(brick_lod, dtype=np.uint8)
    p0 = np.ascontiguousarray(palette0, dtype=np.uint64)
    gl = np.ascontiguousarray(label, dtype=np.uint64)
    gv = np.ascontiguousarray(voxel, dtype=np.uint64)
    gt = np.ascontiguousarray(t_hit, dtype=np.float32)
    cls = np.zeros(w * h, np.uint8)
    el = np.zeros(w * h, np.uint64)
    ev = np.zeros(w * h, np.uint64)
    st = DcStats()
    dc().dc_check_frame(bricks_per_axis, brick_size, empty.ctypes.data_as(c_u8p), lods.ctypes.data_as(c_u8p), ptrs,
                        p0.ctypes.data_as(c_u64p), ctypes.byref(c), w, h, gl.ctypes.data_as(c_u64p),
                        gv.ctypes.data_as(c_u64p), gt.ctypes.data_as(ctypes.POINTER(ctypes.c_float)),
                        threads or (os.cpu_count() or 1), el.ctypes.data_as(c_u64p), ev.ctypes.data_as(c_u64p),
                        cls.ctypes.data_as(c_u8p), ctypes.byref(st))
    stats = {f: getattr(st, f) for f, _ in DcStats._fields_}
    return stats, cls, el, ev
