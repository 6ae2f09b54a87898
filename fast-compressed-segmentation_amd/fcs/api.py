"""Context / Volume / Frame objects over the C ABI."""
import ctypes
import math

import numpy as np

from ._lib import (
    LightingStruct,
    CameraStruct,
    FcsError,
    FrameStruct,
    TimingStruct,
    VolumeInfo,
    c_u8p,
    c_u32p,
    c_u64p,
    c_vp,
    check,
    lib,
)


def device_count():
    return lib.fcs_device_count()


def _ptr(a, typ):
    return a.ctypes.data_as(typ)


class Camera:
    """Camera uniform of camera.js:61-77 plus an orthographic switch."""

    def __init__(self, position, forward, right, up, tan_half_fov=math.tan(math.pi / 6), ortho_half_height=0.0):
        self.s = CameraStruct()
        self.s.position[:] = [float(x) for x in position]
        self.s.forward[:] = [float(x) for x in forward]
        self.s.right[:] = [float(x) for x in right]
        self.s.up[:] = [float(x) for x in up]
        self.s.tan_half_fov = float(tan_half_fov)
        self.s.ortho_half_height = float(ortho_half_height)

    @classmethod
    def ortho_plus_z(cls, grid):
        """Orthographic view along +z covering the whole [0, grid]^2 face."""
        g = float(grid)
        return cls((g / 2, g / 2, -1.0), (0, 0, 1), (1, 0, 0), (0, -1, 0), 0.0, g / 2)

    def as_dict(self):
        s = self.s
        return {
            "position": list(s.position), "forward": list(s.forward), "right": list(s.right), "up": list(s.up),
            "tan_half_fov": s.tan_half_fov, "ortho_half_height": s.ortho_half_height,
        }


class Frame:
    def __init__(self, label, voxel, t_hit, brick_visited, brick_lod_used, bricks_decoded, decode_ms, march_ms,
                 rgba=None):
        self.rgba = rgba
        self.label = label
        self.voxel = voxel
        self.t_hit = t_hit
        self.brick_visited = brick_visited
        self.brick_lod_used = brick_lod_used
        self.bricks_decoded = bricks_decoded
        self.decode_ms = decode_ms
        self.march_ms = march_ms


class Context:
    """One HIP device + one HIP stream (fcs_create)."""

    def __init__(self, device=0):
        h = c_vp()
        rc = lib.fcs_create(device, ctypes.byref(h))
        if rc < 0:
            raise FcsError(rc, "fcs_create failed (no HIP device?)")
        self.h = h

    def close(self):
        if self.h:
            lib.fcs_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def timing(self):
        t = TimingStruct()
        check(lib.fcs_get_timing(self.h, ctypes.byref(t)), self.h)
        return {f: getattr(t, f) for f, _ in TimingStruct._fields_}

    def set_rans_mapping(self, mapping):
        """0 = autotuned, 1 = one lane per brick, 2/3/4 = one wave per brick (scalar-cache table / ballot / LDS
        table), 5/6 = one wave per brick with the byte window in scalar registers (table / ballot)."""
        check(lib.fcs_set_option(self.h, 1, int(mapping)), self.h)

    def set_ops_split(self, split):
        """Op decoder LDS split for bricks up to 32^3: 0 = autotuned per volume,
        1 = three workgroups per CU, 2 = four workgroups per CU."""
        check(lib.fcs_set_option(self.h, 2, int(split)), self.h)

    def set_ops_head(self, mode):
        """Levels up to 3 of each request in a separate one-wave-per-request pass: 0 = for launches of at
        least 1024 requests, 1 = always, 2 = never."""
        check(lib.fcs_set_option(self.h, 3, int(mode)), self.h)

    def set_march_stats(self, on):
        """Host-built frames count what their rays read (timing march_samples,
        march_bricks_entered, march_empty_tested)."""
        check(lib.fcs_set_option(self.h, 4, 1 if on else 0), self.h)

    def sync(self):
        check(lib.fcs_sync(self.h), self.h)

    def load(self, data, brick_size_override=0):
        return Volume(self, data, brick_size_override)


class Volume:
    """A device-resident .csdb volume (fcs_load_csdb)."""

    def __init__(self, ctx, data, brick_size_override=0):
        self.ctx = ctx
        buf = np.frombuffer(bytes(data), dtype=np.uint8)
        h = c_vp()
        check(lib.fcs_load_csdb(ctx.h, _ptr(buf, c_vp), buf.size, brick_size_override, ctypes.byref(h)), ctx.h)
        self.h = h
        self.info = self.get_info()
        self.label_dtype = np.uint64 if self.info["label_bits"] == 64 else np.uint32

    def close(self):
        if self.h:
            lib.fcs_volume_free(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def get_info(self):
        i = VolumeInfo()
        check(lib.fcs_volume_get_info(self.h, ctypes.byref(i)), self.ctx.h)
        d = {f: getattr(i, f) for f, _ in VolumeInfo._fields_}
        d["rans_freq"] = list(i.rans_freq)
        d["rans_cum"] = list(i.rans_cum)
        return d

    def decode_rans(self, raise_on_error=True):
        st = np.zeros(self.info["num_bricks"], np.uint32)
        rc = lib.fcs_decode_rans(self.h, _ptr(st, c_u32p))
        if rc < 0 and (raise_on_error or rc != -6):
            check(rc, self.ctx.h)
        return st

    def cache_reserve(self, voxels):
        check(lib.fcs_cache_reserve(self.h, int(voxels)), self.ctx.h)

    def cache_reserve_direct(self):
        """Direct-mapped pool: brick i owns slot i (fcs_cache_reserve_direct)."""
        check(lib.fcs_cache_reserve_direct(self.h), self.ctx.h)

    def cache_resident(self):
        """fcs_cache_resident: (resident bricks ascending, the LOD each holds)."""
        n = check(lib.fcs_cache_resident(self.h, None, None, 0), self.ctx.h)
        b = np.zeros(n, np.uint32)
        t = np.zeros(n, np.uint8)
        check(lib.fcs_cache_resident(self.h, _ptr(b, c_u32p), _ptr(t, c_u8p), n), self.ctx.h)
        return b, t

    def cache_clear(self):
        check(lib.fcs_cache_clear(self.h), self.ctx.h)

    def request_bricks(self, bricks, lods, raise_on_error=True):
        b = np.ascontiguousarray(bricks, dtype=np.uint32)
        l = np.ascontiguousarray(lods, dtype=np.uint8)
        off = np.zeros(b.size, np.uint64)
        st = np.zeros(b.size, np.uint32)
        rc = lib.fcs_request_bricks(self.h, _ptr(b, c_u32p), _ptr(l, c_u8p), b.size, _ptr(off, c_u64p),
                                    _ptr(st, c_u32p))
        if rc < 0 and (raise_on_error or rc != -6):
            check(rc, self.ctx.h)
        return off, st

    def request_bricks_static(self, bricks, lods, raise_on_error=True):
        """fcs_request_bricks_static: the same as request_bricks with no device
        allocation (the cache pool must be reserved and decode_rans done)."""
        b = np.ascontiguousarray(bricks, dtype=np.uint32)
        l = np.ascontiguousarray(lods, dtype=np.uint8)
        off = np.zeros(b.size, np.uint64)
        st = np.zeros(b.size, np.uint32)
        rc = lib.fcs_request_bricks_static(self.h, _ptr(b, c_u32p), _ptr(l, c_u8p), b.size, _ptr(off, c_u64p),
                                           _ptr(st, c_u32p))
        if rc < 0 and (raise_on_error or rc != -6):
            check(rc, self.ctx.h)
        return off, st

    def decode_all(self, lod=255, pyramid=True):
        check(lib.fcs_decode_all(self.h, int(lod), 1 if pyramid else 0), self.ctx.h)

    def decode_volume(self, lod=255, pyramid=True, raise_on_error=True):
        """decode_rans() then decode_all(lod, pyramid) in one call, both launches
        back to back on the device (fcs_decode_volume). Returns per-brick status
        words of both stages; the context timing's pipeline_ms covers the whole
        call."""
        st = np.zeros(self.info["num_bricks"], np.uint32)
        rc = lib.fcs_decode_volume(self.h, int(lod), 1 if pyramid else 0, _ptr(st, c_u32p))
        if rc < 0 and (raise_on_error or rc != -6):
            check(rc, self.ctx.h)
        return st

    def read_cache(self, offset, count):
        out = np.zeros(int(count), self.label_dtype)
        check(lib.fcs_read_cache(self.h, int(offset), int(count), _ptr(out, c_vp)), self.ctx.h)
        return out

    def read_brick(self, brick, lod):
        out = np.zeros(1 << (3 * lod), self.label_dtype)
        check(lib.fcs_read_brick(self.h, int(brick), int(lod), _ptr(out, c_vp)), self.ctx.h)
        return out

    def read_nibbles(self, brick):
        """Nibble stream of a brick after decode_rans, two nibbles per byte,
        first in the high half (decodeAllRansBricks, web/scripts/decode.js:17-26)."""
        n = check(lib.fcs_read_nibbles(self.h, int(brick), None, 0), self.ctx.h)
        out = np.zeros(n, np.uint8)
        check(lib.fcs_read_nibbles(self.h, int(brick), _ptr(out, c_u8p), n), self.ctx.h)
        return out

    def brick_lods(self, cam, w, h):
        out = np.zeros(self.info["num_bricks"], np.uint8)
        check(lib.fcs_brick_lods(self.h, ctypes.byref(cam.s), w, h, _ptr(out, c_u8p)), self.ctx.h)
        return out

    def orbit_camera(self, yaw, pitch, radius, fov=math.pi / 3):
        cam = Camera((0, 0, 0), (0, 0, 1), (1, 0, 0), (0, 1, 0))
        check(lib.fcs_camera_orbit(self.h, yaw, pitch, radius, fov, ctypes.byref(cam.s)), self.ctx.h)
        return cam

    def render_visible(self, cam, w, h, brick_lod=None, want=("brick_visited",)):
        """Frame decoding only the bricks its rays enter (fcs_render_visible)."""
        fr, f = self._frame_buffers(w, h, want)
        lp = None
        if brick_lod is not None:
            bl = np.ascontiguousarray(brick_lod, dtype=np.uint8)
            lp = _ptr(bl, c_u8p)
        check(lib.fcs_render_visible(self.h, ctypes.byref(cam.s), lp, w, h, ctypes.byref(f)), self.ctx.h)
        return self._frame(fr, f)

    def prefetch(self, bricks, lods):
        """Decode (brick, LOD) pairs ahead, asynchronously (fcs_prefetch_bricks)."""
        b = np.ascontiguousarray(bricks, dtype=np.uint32)
        l = np.ascontiguousarray(lods, dtype=np.uint8)
        return check(lib.fcs_prefetch_bricks(self.h, _ptr(b, c_u32p), _ptr(l, c_u8p), b.size), self.ctx.h)

    def frame_checksum(self):
        out = ctypes.c_uint64()
        check(lib.fcs_frame_checksum(self.h, ctypes.byref(out)), self.ctx.h)
        return out.value

    def camera_features(self, cam):
        out = (ctypes.c_float * 12)()
        check(lib.fcs_camera_features(self.h, ctypes.byref(cam.s), out), self.ctx.h)
        return np.array(out[:], np.float32)

    def render_prefetch(self, mlp, cam, w, h, cam_ahead=None, threshold=0.5, max_prefetch=4096,
                        want=("brick_visited",)):
        """One frame with learned prefetch (fcs_render_prefetch); returns (Frame, bricks prefetched)."""
        fr, f = self._frame_buffers(w, h, want)
        n = ctypes.c_uint32()
        lst = np.full(max(1, max_prefetch), 0xFFFFFFFF, np.uint32)
        ca = ctypes.byref(cam_ahead.s) if cam_ahead is not None else None
        check(lib.fcs_render_prefetch(self.h, mlp.h, ctypes.byref(cam.s), ca, w, h, threshold, max_prefetch,
                                      ctypes.byref(f), ctypes.byref(n), _ptr(lst, c_u32p)), self.ctx.h)
        return self._frame(fr, f), lst[lst != 0xFFFFFFFF]

    def _frame_buffers(self, w, h, want):
        n = w * h
        nb = self.info["num_bricks"]
        fr = {
            "label": np.zeros(n, self.label_dtype) if "label" in want else None,
            "voxel": np.zeros(n, np.uint64) if "voxel" in want else None,
            "t_hit": np.zeros(n, np.float32) if "t_hit" in want else None,
            "vis": np.zeros(nb, np.uint32) if "brick_visited" in want else None,
            "used": np.zeros(nb, np.uint8),
            "rgba": np.zeros(n, np.uint32) if "rgba" in want else None,
        }
        f = FrameStruct()
        f.rgba = _ptr(fr["rgba"], c_u32p) if fr["rgba"] is not None else None
        f.label = fr["label"].ctypes.data if fr["label"] is not None else None
        f.voxel = _ptr(fr["voxel"], c_u64p) if fr["voxel"] is not None else None
        f.t_hit = fr["t_hit"].ctypes.data_as(ctypes.POINTER(ctypes.c_float)) if fr["t_hit"] is not None else None
        f.brick_visited = _ptr(fr["vis"], c_u32p) if fr["vis"] is not None else None
        f.brick_lod_used = _ptr(fr["used"], c_u8p)
        return fr, f

    @staticmethod
    def _frame(fr, f):
        return Frame(fr["label"], fr["voxel"], fr["t_hit"], fr["vis"], fr["used"], f.bricks_decoded, f.decode_ms,
                     f.march_ms, fr["rgba"])

    def render(self, cam, w, h, brick_lod=None, want=("label", "voxel", "t_hit", "brick_visited")):
        """fcs_render; `want` names the outputs to copy back ("label", "voxel",
        "t_hit", "brick_visited", "rgba": the shaded composite)."""
        fr, f = self._frame_buffers(w, h, want)
        lp = None
        if brick_lod is not None:
            bl = np.ascontiguousarray(brick_lod, dtype=np.uint8)
            lp = _ptr(bl, c_u8p)
        check(lib.fcs_render(self.h, ctypes.byref(cam.s), lp, w, h, ctypes.byref(f)), self.ctx.h)
        return self._frame(fr, f)

    def set_lighting(self, light_angle_deg=25.0, ambient=0.5, diffuse_strength=0.8):
        """updateLightingOptions (webgpuSetup.js:186-222)."""
        ls = LightingStruct(light_angle_deg, ambient, diffuse_strength, 0.0)
        check(lib.fcs_set_lighting(self.h, ctypes.byref(ls)), self.ctx.h)

    def set_label_visibility(self, hidden=(), highlight=None):
        """applyLabelVisibility + applyLabelHover (webgpuSetup.js:827-882)."""
        hs = np.ascontiguousarray(np.array(list(hidden), dtype=np.uint64))
        hl = ctypes.c_uint64(int(highlight)) if highlight is not None else None
        check(lib.fcs_set_label_visibility(self.h, _ptr(hs, c_u64p) if hs.size else None, hs.size,
                                           ctypes.byref(hl) if hl is not None else None), self.ctx.h)

    def transfer_function(self):
        """getDatasetLabels + the current colours: (labels ascending, RGBA8)."""
        n = check(lib.fcs_transfer_function(self.h, None, None, 0), self.ctx.h)
        labels = np.zeros(n, np.uint64)
        rgba = np.zeros(n, np.uint32)
        check(lib.fcs_transfer_function(self.h, _ptr(labels, c_u64p), _ptr(rgba, c_u32p), n), self.ctx.h)
        return labels, rgba

    def lighting_direction(self, light_angle_deg=25.0):
        """The light direction of the lighting uniform, float32."""
        a = light_angle_deg * math.pi / 180.0
        return np.array([math.sin(a), 0.6, -math.cos(a)], np.float32)


def load_dataset_from_array_buffer(ctx, array_buffer, brick_size_override=0):
    """loadDatasetFromArrayBuffer (web/scripts/datasetLoader.js:39-145):
    parse a .csdb byte buffer; raises FcsError("Invalid file format") on a bad
    magic, like the reference."""
    return Volume(ctx, array_buffer, brick_size_override)


def decode_all_rans_bricks(volume):
    """decodeAllRansBricks (web/scripts/decode.js:17-26), on the device: every
    brick's rANS stream becomes a device-resident nibble stream. Returns the
    per-brick status words."""
    return volume.decode_rans()
