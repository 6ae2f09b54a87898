"""fcs_request_bricks_static (the allocation-free request entry point) against
fcs_request_bricks on the same inputs: offsets, status words and every label
of every request equal, on two calls in a row, the second with fewer
requests; requests beyond its maximum and calls without a reserved pool are
rejected."""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def pair(gpu_ctx):
    import fcs

    print("[no-alloc] encoding a 512^3 volume with 32^3 bricks", flush=True)
    data = fcs.encode_generated(fcs.Generator(5, size=512), 32)
    va, vs = gpu_ctx.load(data), gpu_ctx.load(data)
    for v in (va, vs):
        assert not np.any(v.decode_rans())
        v.cache_reserve(v.info["num_bricks"] * 32768)
    yield va, vs
    va.close()
    vs.close()


def _regions(v, off, lods):
    sizes = np.left_shift(np.uint64(1), np.uint64(3) * lods.astype(np.uint64))
    whole = v.read_cache(0, int((off + sizes).max()))
    return [whole[int(o):int(o + n)] for o, n in zip(off, sizes)]


def test_static_requests_equal_allocating_requests(pair):
    va, vs = pair
    nb = va.info["num_bricks"]
    rng = np.random.default_rng(7)
    for n in (nb, 700):  # the second call smaller than the first
        bricks = rng.choice(nb, n, replace=False).astype(np.uint32)
        lods = rng.integers(0, 6, n).astype(np.uint8)
        for v in (va, vs):
            v.cache_clear()
        oa, sa = va.request_bricks(bricks, lods)
        print(f"[no-alloc] {n} requests through the allocating entry point done", flush=True)
        os_, ss = vs.request_bricks_static(bricks, lods)
        print(f"[no-alloc] {n} requests through the static entry point done", flush=True)
        assert np.array_equal(oa, os_) and np.array_equal(sa, ss)
        assert not np.any(ss)
        ra, rs = _regions(va, oa, lods), _regions(vs, os_, lods)
        bad = [k for k in range(n) if not np.array_equal(ra[k], rs[k])]
        assert not bad, bad[:8]


def test_static_requests_reject_what_they_cannot_do(gpu_ctx, pair):
    import fcs
    from fcs._lib import STATIC_MAX_REQUESTS, lib

    va, _ = pair
    n = STATIC_MAX_REQUESTS + 1
    b = np.zeros(n, np.uint32)
    t = np.zeros(n, np.uint8)
    off = np.zeros(n, np.uint64)
    st = np.zeros(n, np.uint32)
    rc = lib.fcs_request_bricks_static(va.h, b.ctypes.data_as(ctypes.POINTER(ctypes.c_uint32)),
                                       t.ctypes.data_as(ctypes.POINTER(ctypes.c_uint8)), n,
                                       off.ctypes.data_as(ctypes.POINTER(ctypes.c_uint64)),
                                       st.ctypes.data_as(ctypes.POINTER(ctypes.c_uint32)))
    assert rc == -1  # FCS_E_INVALID
    # a fresh volume has no pool and no rANS stage: rejected, nothing allocated
    from conftest import golden_bytes

    fresh = gpu_ctx.load(golden_bytes("cfg1_b32"), 32)
    try:
        with pytest.raises(fcs.FcsError) as e:
            fresh.request_bricks_static([0], [5])
        assert e.value.code == -1
    finally:
        fresh.close()
