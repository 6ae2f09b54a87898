"""Input checks of the C ABI that need a context (a HIP device): failed
requests' offsets and oversize regions are rejected by fcs_read_cache, and a
header whose brick count the input cannot hold is rejected before anything
is sized by it."""
import ctypes
import struct

import numpy as np
import pytest

from conftest import golden_bytes

pytestmark = pytest.mark.gpu


def test_read_cache_rejects_failed_offsets_and_wrapping_regions(gpu_ctx):
    import fcs
    from fcs._lib import lib

    vol = gpu_ctx.load(golden_bytes("cfg1_b32"), 32)
    try:
        vol.cache_reserve(32768)  # room for one LOD-5 brick
        # a request the pool cannot hold fails with offset ~0
        off, st = vol.request_bricks([0, 1], [5, 5], raise_on_error=False)
        assert st[0] == 0 and off[1] == np.uint64(2 ** 64 - 1) and st[1] & 8  # FCS_ST_CACHE_FULL
        buf = np.zeros(16, np.uint32)
        p = buf.ctypes.data_as(ctypes.c_void_p)
        assert lib.fcs_read_cache(vol.h, ctypes.c_uint64(int(off[1])), 1, p) == -1
        assert lib.fcs_read_cache(vol.h, 0, ctypes.c_uint64(2 ** 64 - 1), p) == -1
        assert lib.fcs_read_cache(vol.h, 32768 - 16, 17, p) == -1
        assert lib.fcs_read_cache(vol.h, 32768 - 16, 16, p) == 0
    finally:
        vol.close()


def test_load_rejects_brick_count_beyond_the_stream(gpu_ctx):
    import fcs

    head = bytearray(golden_bytes("cfg1_b32")[:152])
    head[8:12] = struct.pack("<I", 1 << 30)  # 1024^3 bricks in a 152-byte stream
    with pytest.raises(fcs.FcsError) as e:
        gpu_ctx.load(bytes(head) + b"\0" * 64, 32)
    assert e.value.code == -2  # FCS_E_FORMAT
