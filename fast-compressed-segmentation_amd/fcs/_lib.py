"""ctypes loader for libfcs.so. Fails loudly when the native library is absent."""
import ctypes
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
lib_path = os.path.join(os.path.dirname(_HERE), "lib", "libfcs.so")

ST_STREAM_OVERRUN = 1
ST_PALETTE_RANGE = 2
ST_NSYM_MISMATCH = 4
ST_CACHE_FULL = 8
ST_RANS_STATE = 16
ST_BAD_OP = 32
ST_BAD_REQUEST = 64

E_NAMES = {
    -1: "FCS_E_INVALID",
    -2: "FCS_E_FORMAT",
    -3: "FCS_E_HIP",
    -4: "FCS_E_NOMEM",
    -5: "FCS_E_NODEVICE",
    -6: "FCS_E_CORRUPT",
    -7: "FCS_E_CACHE_FULL",
}


class FcsError(RuntimeError):
    def __init__(self, code, msg=""):
        self.code = code
        super().__init__(f"{E_NAMES.get(code, code)}: {msg}")


if not os.path.exists(lib_path):
    raise ImportError(
        f"libfcs.so not found at {lib_path}; build it with "
        "`python fast-compressed-segmentation_amd/build.py` (or __graft_entry__.build())"
    )
lib = ctypes.CDLL(lib_path)

c_u32p = ctypes.POINTER(ctypes.c_uint32)
c_u64p = ctypes.POINTER(ctypes.c_uint64)
c_u8p = ctypes.POINTER(ctypes.c_uint8)
c_vp = ctypes.c_void_p


class VolumeInfo(ctypes.Structure):
    _fields_ = [
        ("num_bricks", ctypes.c_uint32),
        ("brick_size", ctypes.c_uint32),
        ("levels", ctypes.c_uint32),
        ("bricks_per_axis", ctypes.c_uint32),
        ("grid_size", ctypes.c_uint32),
        ("label_bits", ctypes.c_uint32),
        ("header_brick_size", ctypes.c_uint32),
        ("rans_ready", ctypes.c_uint32),
        ("compressed_bytes", ctypes.c_uint64),
        ("total_symbols", ctypes.c_uint64),
        ("palette_entries", ctypes.c_uint64),
        ("cache_capacity_voxels", ctypes.c_uint64),
        ("rans_freq", ctypes.c_uint32 * 16),
        ("rans_cum", ctypes.c_uint32 * 16),
    ]


class CameraStruct(ctypes.Structure):
    _fields_ = [
        ("position", ctypes.c_float * 3),
        ("tan_half_fov", ctypes.c_float),
        ("forward", ctypes.c_float * 3),
        ("ortho_half_height", ctypes.c_float),
        ("right", ctypes.c_float * 3),
        ("pad1", ctypes.c_float),
        ("up", ctypes.c_float * 3),
        ("pad2", ctypes.c_float),
    ]


class FrameStruct(ctypes.Structure):
    _fields_ = [
        ("label", c_vp),
        ("voxel", c_u64p),
        ("t_hit", ctypes.POINTER(ctypes.c_float)),
        ("brick_visited", c_u32p),
        ("brick_lod_used", c_u8p),
        ("bricks_decoded", ctypes.c_uint32),
        ("pad", ctypes.c_uint32),
        ("decode_ms", ctypes.c_float),
        ("march_ms", ctypes.c_float),
        ("rgba", c_u32p),
    ]


class LightingStruct(ctypes.Structure):
    _fields_ = [("light_angle_deg", ctypes.c_float), ("ambient", ctypes.c_float),
                ("diffuse_strength", ctypes.c_float), ("pad", ctypes.c_float)]


class TimingStruct(ctypes.Structure):
    _fields_ = [
        ("rans_ms", ctypes.c_float),
        ("decode_ms", ctypes.c_float),
        ("march_ms", ctypes.c_float),
        ("rans_mapping", ctypes.c_uint32),
        ("decoded_voxels", ctypes.c_uint64),
        ("decode_bytes", ctypes.c_uint64),
        ("rans_symbols", ctypes.c_uint64),
        ("rans_bytes", ctypes.c_uint64),
        ("tune_lane_ms", ctypes.c_float),
        ("tune_wave_ms", ctypes.c_float),
        ("tune_ballot_ms", ctypes.c_float),
        ("tune_wave_lds_ms", ctypes.c_float),
        ("ops_split", ctypes.c_uint32),
        ("tune_ops_3wg_ms", ctypes.c_float),
        ("tune_ops_4wg_ms", ctypes.c_float),
        ("tune_win_ms", ctypes.c_float),
        ("tune_win_ballot_ms", ctypes.c_float),
        ("pipeline_ms", ctypes.c_float),
        ("march_samples", ctypes.c_uint64),
        ("march_bricks_entered", ctypes.c_uint64),
        ("march_empty_tested", ctypes.c_uint64),
    ]


RANS_AUTO, RANS_LANE, RANS_WAVE, RANS_WAVE_BALLOT, RANS_WAVE_LDS, RANS_WAVE_WIN, RANS_WAVE_WIN_BALLOT = \
    0, 1, 2, 3, 4, 5, 6
OPT_RANS_MAPPING = 1
OPT_OPS_SPLIT = 2
OPS_SPLIT_AUTO, OPS_SPLIT_3WG, OPS_SPLIT_4WG = 0, 1, 2
OPT_OPS_HEAD = 3
OPS_HEAD_AUTO, OPS_HEAD_ON, OPS_HEAD_OFF = 0, 1, 2


class GenParams(ctypes.Structure):
    _fields_ = [
        ("kind", ctypes.c_uint32),
        ("size", ctypes.c_uint32),
        ("seed", ctypes.c_uint64),
        ("num_seeds", ctypes.c_uint32),
        ("wx", ctypes.c_uint32),
        ("wy", ctypes.c_uint32),
        ("wz", ctypes.c_uint32),
        ("noise_amp", ctypes.c_uint32),
        ("background", ctypes.c_uint32),
    ]


class EncodeOpts(ctypes.Structure):
    _fields_ = [("brick_size", ctypes.c_uint32), ("header_brick_size_64", ctypes.c_uint32), ("threads", ctypes.c_int),
                ("progress", ctypes.c_uint32)]


def _sig(name, res, *args):
    f = getattr(lib, name)
    f.restype = res
    f.argtypes = list(args)
    return f


_sig("fcs_device_count", ctypes.c_int)
_sig("fcs_create", ctypes.c_int, ctypes.c_int, ctypes.POINTER(c_vp))
_sig("fcs_destroy", None, c_vp)
_sig("fcs_last_error", ctypes.c_char_p, c_vp)
_sig("fcs_sync", ctypes.c_int, c_vp)
_sig("fcs_get_timing", ctypes.c_int, c_vp, ctypes.POINTER(TimingStruct))
_sig("fcs_set_option", ctypes.c_int, c_vp, ctypes.c_int, ctypes.c_int64)
_sig("fcs_load_csdb", ctypes.c_int, c_vp, c_vp, ctypes.c_size_t, ctypes.c_uint32, ctypes.POINTER(c_vp))
_sig("fcs_volume_free", None, c_vp)
_sig("fcs_volume_get_info", ctypes.c_int, c_vp, ctypes.POINTER(VolumeInfo))
_sig("fcs_decode_rans", ctypes.c_int, c_vp, c_u32p)
_sig("fcs_cache_reserve", ctypes.c_int, c_vp, ctypes.c_uint64)
_sig("fcs_cache_clear", ctypes.c_int, c_vp)
_sig("fcs_cache_reserve_direct", ctypes.c_int, c_vp)
_sig("fcs_request_bricks", ctypes.c_int, c_vp, c_u32p, c_u8p, ctypes.c_uint32, c_u64p, c_u32p)
_sig("fcs_cache_resident", ctypes.c_int64, c_vp, c_u32p, c_u8p, ctypes.c_uint64)
_sig("fcs_request_bricks_static", ctypes.c_int, c_vp, c_u32p, c_u8p, ctypes.c_uint32, c_u64p, c_u32p)
STATIC_MAX_REQUESTS = 262144
_sig("fcs_decode_all", ctypes.c_int, c_vp, ctypes.c_uint32, ctypes.c_int)
_sig("fcs_decode_volume", ctypes.c_int, c_vp, ctypes.c_uint32, ctypes.c_int, c_u32p)
_sig("fcs_read_brick", ctypes.c_int, c_vp, ctypes.c_uint32, ctypes.c_uint32, c_vp)
_sig("fcs_read_cache", ctypes.c_int, c_vp, ctypes.c_uint64, ctypes.c_uint64, c_vp)
_sig("fcs_read_nibbles", ctypes.c_int64, c_vp, ctypes.c_uint32, c_u8p, ctypes.c_uint64)
_sig("fcs_brick_lods", ctypes.c_int, c_vp, ctypes.POINTER(CameraStruct), ctypes.c_uint32, ctypes.c_uint32, c_u8p)
_sig("fcs_render", ctypes.c_int, c_vp, ctypes.POINTER(CameraStruct), c_u8p, ctypes.c_uint32, ctypes.c_uint32,
     ctypes.POINTER(FrameStruct))
_sig("fcs_camera_orbit", ctypes.c_int, c_vp, ctypes.c_double, ctypes.c_double, ctypes.c_double, ctypes.c_double,
     ctypes.POINTER(CameraStruct))
_sig("fcs_render_visible", ctypes.c_int, c_vp, ctypes.POINTER(CameraStruct), c_u8p, ctypes.c_uint32, ctypes.c_uint32,
     ctypes.POINTER(FrameStruct))
_sig("fcs_prefetch_bricks", ctypes.c_int, c_vp, c_u32p, c_u8p, ctypes.c_uint32)
_sig("fcs_frame_checksum", ctypes.c_int, c_vp, c_u64p)
_sig("fcs_camera_features", ctypes.c_int, c_vp, ctypes.POINTER(CameraStruct), ctypes.POINTER(ctypes.c_float))
_sig("fcs_render_prefetch", ctypes.c_int, c_vp, c_vp, ctypes.POINTER(CameraStruct), ctypes.POINTER(CameraStruct),
     ctypes.c_uint32, ctypes.c_uint32, ctypes.c_float, ctypes.c_uint32, ctypes.POINTER(FrameStruct),
     ctypes.POINTER(ctypes.c_uint32), c_u32p)
_sig("fcs_gen_preset", ctypes.c_int, ctypes.c_int, ctypes.c_uint32, ctypes.POINTER(GenParams))
_sig("fcs_gen_create", ctypes.c_int, ctypes.POINTER(GenParams), ctypes.POINTER(c_vp))
_sig("fcs_gen_destroy", None, c_vp)
_sig("fcs_gen_box", ctypes.c_int, c_vp, ctypes.c_uint32, ctypes.c_uint32, ctypes.c_uint32, ctypes.c_uint32,
     ctypes.c_uint32, ctypes.c_uint32, c_u32p, ctypes.c_int)
_sig("fcs_gen_bricks", ctypes.c_int, c_vp, ctypes.c_uint32, ctypes.c_uint32, ctypes.c_uint32, c_u32p, ctypes.c_int)
_sig("fcs_gen_background_fraction", ctypes.c_double, c_vp)
_sig("fcs_encode_volume", ctypes.c_int, c_vp, ctypes.c_uint32, ctypes.c_uint32, ctypes.c_uint32, ctypes.c_uint32,
     ctypes.POINTER(EncodeOpts), ctypes.POINTER(c_u8p), ctypes.POINTER(ctypes.c_size_t))
_sig("fcs_encode_generated", ctypes.c_int, c_vp, ctypes.POINTER(EncodeOpts), ctypes.POINTER(c_u8p),
     ctypes.POINTER(ctypes.c_size_t))
_sig("fcs_free_buffer", None, c_vp)
_sig("fcs_brick_pyramid", ctypes.c_int, c_u64p, ctypes.c_uint32, c_u64p, c_u8p)


class ViewResultStruct(ctypes.Structure):
    _fields_ = [
        ("checksum", ctypes.c_uint64),
        ("misses", ctypes.c_uint32),
        ("w", ctypes.c_uint32),
        ("h", ctypes.c_uint32),
        ("label_bits", ctypes.c_uint32),
        ("device_ms", ctypes.c_float),
        ("pad", ctypes.c_uint32),
    ]


class ServeRequestStruct(ctypes.Structure):
    _fields_ = [
        ("arrival_s", ctypes.c_double),
        ("session", ctypes.c_uint32),
        ("volume", ctypes.c_uint32),
        ("w", ctypes.c_uint32),
        ("h", ctypes.c_uint32),
        ("cam", CameraStruct),
    ]


class ServeResultStruct(ctypes.Structure):
    _fields_ = [
        ("issue_s", ctypes.c_double),
        ("done_s", ctypes.c_double),
        ("checksum", ctypes.c_uint64),
        ("device_ms", ctypes.c_float),
        ("misses", ctypes.c_uint32),
        ("view", ctypes.c_uint32),
        ("pad", ctypes.c_uint32),
    ]


VIEW_GRAPH, VIEW_LABELS, VIEW_CHECKSUM, VIEW_RGBA = 1, 2, 4, 8

_sig("fcs_view_create", ctypes.c_int, c_vp, ctypes.c_uint32, ctypes.c_uint32, ctypes.c_uint32, ctypes.POINTER(c_vp))
_sig("fcs_view_destroy", None, c_vp)
_sig("fcs_view_render_async", ctypes.c_int, c_vp, c_vp, ctypes.POINTER(CameraStruct), c_u8p, ctypes.c_uint32,
     ctypes.c_uint32)
_sig("fcs_view_query", ctypes.c_int, c_vp)
_sig("fcs_view_wait", ctypes.c_int, c_vp, ctypes.POINTER(ViewResultStruct))
_sig("fcs_view_labels", c_vp, c_vp)
_sig("fcs_view_read", ctypes.c_int, c_vp, c_vp, c_u64p, ctypes.POINTER(ctypes.c_float), c_u8p)
_sig("fcs_view_graphs", ctypes.c_uint32, c_vp)
_sig("fcs_view_read_rgba", ctypes.c_int, c_vp, c_u32p)
_sig("fcs_serve_plan", ctypes.c_int, c_u64p, ctypes.POINTER(ctypes.c_double), ctypes.c_uint32, ctypes.c_uint64,
     c_u32p, c_u64p)
_sig("fcs_serve_admit", ctypes.c_int, ctypes.POINTER(c_vp), ctypes.POINTER(ctypes.c_double), ctypes.c_uint32,
     ctypes.c_uint64, c_u32p, c_u64p)
_sig("fcs_serve_run", ctypes.c_int, ctypes.POINTER(c_vp), ctypes.c_uint32, ctypes.POINTER(c_vp), ctypes.c_uint32,
     ctypes.POINTER(ServeRequestStruct), ctypes.c_uint32, ctypes.POINTER(ServeResultStruct))


_sig("fcs_set_lighting", ctypes.c_int, c_vp, ctypes.POINTER(LightingStruct))
_sig("fcs_set_label_visibility", ctypes.c_int, c_vp, c_u64p, ctypes.c_uint32, c_u64p)
_sig("fcs_transfer_function", ctypes.c_int64, c_vp, c_u64p, c_u32p, ctypes.c_uint64)
_sig("fcs_tf_default_rgba", ctypes.c_uint32, ctypes.c_uint64)


def check(rc, ctx=None):
    if rc < 0:
        msg = lib.fcs_last_error(ctx).decode() if ctx else ""
        raise FcsError(rc, msg)
    return rc
