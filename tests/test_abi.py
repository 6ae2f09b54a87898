"""The C-ABI library loads and exports every function include/fcs/*.h declares."""
import ctypes
import os
import re

import pytest

from conftest import ROOT

HEADERS = [os.path.join(ROOT, "include", "fcs", h) for h in ("fcs.h", "fcs_producer.h", "fcs_mlp.h",
                                                                  "fcs_serve.h")]


def declared_functions():
    names = set()
    for h in HEADERS:
        src = open(h).read()
        src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
        for m in re.finditer(r"^[A-Za-z_][\w \*]*?\b(fcs_\w+)\s*\(", src, flags=re.M):
            names.add(m.group(1))
    return sorted(names)


def test_header_declares_the_boundary():
    names = declared_functions()
    for required in ("fcs_create", "fcs_destroy", "fcs_load_csdb", "fcs_volume_get_info", "fcs_decode_rans",
                     "fcs_request_bricks", "fcs_decode_all", "fcs_read_brick", "fcs_render", "fcs_last_error",
                     "fcs_brick_lods", "fcs_camera_orbit", "fcs_encode_volume", "fcs_gen_create"):
        assert required in names


def test_library_exports_every_declared_symbol():
    import fcs

    lib = ctypes.CDLL(fcs.lib_path)
    missing = [n for n in declared_functions() if not hasattr(lib, n)]
    assert not missing, missing


def test_no_device_is_reported_not_faked():
    import fcs

    if fcs.device_count() > 0:
        pytest.skip("a HIP device is present")
    with pytest.raises(fcs.FcsError) as e:
        fcs.Context(0)
    assert e.value.code == -5  # FCS_E_NODEVICE


def test_decode_entry_points_reject_a_null_volume():
    """The decode entry points check their volume before any device call."""
    import fcs

    lib = ctypes.CDLL(fcs.lib_path)
    for name, args in (("fcs_decode_rans", (None, None)), ("fcs_decode_all", (None, 5, 1)),
                       ("fcs_decode_volume", (None, 5, 1, None))):
        fn = getattr(lib, name)
        fn.restype = ctypes.c_int
        assert fn(*args) == -1, name  # FCS_E_INVALID
