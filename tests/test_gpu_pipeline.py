"""fcs_decode_volume: the rANS stage and the full decode in one call, both
launches back to back on the context stream. Bit-exact against decode_rans +
decode_all (and the oracle): pyramids and single levels, both rANS mappings,
both label widths, corrupt streams, and a 512^3 volume of 64^3 bricks."""
import math

import numpy as np
import pytest

import fcs
from conftest import golden_bytes
from oracle import oracle as O

pytestmark = pytest.mark.gpu

CASES = ["cfg1_b32", "rand_b4", "rand_b8", "const_mix_b8", "checker_b16", "u64_b8", "microns8_b64"]


@pytest.fixture(scope="module")
def ctx():
    if fcs.device_count() == 0:
        pytest.skip("no HIP device")
    c = fcs.Context(0)
    yield c
    c.close()


def _full(ctx, data, b, pipelined, lod=255, pyramid=True, mapping=0):
    ctx.set_rans_mapping(mapping)
    vol = ctx.load(data, b)
    try:
        L, nb = vol.info["levels"], vol.info["num_bricks"]
        t = min(lod, L)
        per = O.pyramid_offset(t + 1) if pyramid else 1 << (3 * t)
        if pipelined:
            st = vol.decode_volume(lod, pyramid)
            assert ctx.timing()["pipeline_ms"] > 0
        else:
            st = vol.decode_rans()
            vol.decode_all(lod, pyramid)
        return st.copy(), vol.read_cache(0, per * nb).reshape(nb, per).copy()
    finally:
        vol.close()
        ctx.set_rans_mapping(0)


@pytest.mark.parametrize("name", CASES)
def test_decode_volume_pyramids_match_two_calls_and_oracle(ctx, manifest, name):
    print(f"[gpu] decode_volume pyramids {name}", flush=True)
    b = manifest[name]["brick_size"]
    data = golden_bytes(name)
    st, got = _full(ctx, data, b, True)
    _, want = _full(ctx, data, b, False)
    assert np.all(st == 0)
    assert np.array_equal(got, want)
    ds = O.Dataset(data, b)
    for i in range(0, ds.num_bricks, max(1, ds.num_bricks // 6)):
        for t in range(ds.L + 1):
            lab, s = ds.decode(i, t)
            assert s == 0
            seg = got[i, O.pyramid_offset(t):O.pyramid_offset(t + 1)].astype(np.uint64)
            assert np.array_equal(seg, lab), (name, i, t)


@pytest.mark.parametrize("mapping", [1, 2])
@pytest.mark.parametrize("name", ["rand_b8", "microns8_b64"])
def test_decode_volume_forced_mappings(ctx, manifest, name, mapping):
    print(f"[gpu] decode_volume {name} mapping {mapping}", flush=True)
    b = manifest[name]["brick_size"]
    data = golden_bytes(name)
    L = O.Dataset(data, b).L
    for lod, pyr in ((L, False), (max(L - 1, 0), True)):
        st, got = _full(ctx, data, b, True, lod, pyr, mapping)
        _, want = _full(ctx, data, b, False, lod, pyr, mapping)
        assert np.all(st == 0)
        assert np.array_equal(got, want), (name, mapping, lod, pyr)


def test_decode_volume_corrupt_streams(ctx):
    """Damaged streams: per-brick statuses are the rANS stage's bits or'ed with
    the op decoder's, and every brick's labels equal the two-call path's; a
    second call (which launches the requests the first one staged) reports
    the same statuses and labels."""
    print("[gpu] decode_volume corrupt streams", flush=True)
    data = bytearray(golden_bytes("rand_b8"))
    ds = O.Dataset(bytes(data), 8)
    for i in range(0, ds.num_bricks, 3):
        br = ds.brick(i)
        if br.E > 8:
            data[br.str_off + br.E // 3] ^= 0x5A
    data = bytes(data)
    L, nb = ds.L, ds.num_bricks
    n = 1 << (3 * L)
    vol = ctx.load(data, 8)
    st = vol.decode_volume(L, False, raise_on_error=False)
    got = vol.read_cache(0, n * nb).reshape(nb, n).copy()
    st2 = vol.decode_volume(L, False, raise_on_error=False)
    assert np.array_equal(st2, st)
    assert np.array_equal(vol.read_cache(0, n * nb).reshape(nb, n), got)
    vol.close()
    vol = ctx.load(data, 8)
    rst = vol.decode_rans(raise_on_error=False)
    vol.cache_reserve(nb * n)
    off, ost = vol.request_bricks(np.arange(nb, dtype=np.uint32), np.full(nb, L, np.uint8), raise_on_error=False)
    want = np.stack([vol.read_cache(int(o), n) for o in off])
    vol.close()
    assert np.any(st != 0)
    assert np.array_equal(st, rst | ost)
    assert np.array_equal(got, want)


def test_decode_volume_512cube_b64_splits(ctx):
    """512^3 Voronoi volume, 64^3 bricks (512 streams of very different
    lengths): the pipelined pyramid equals the two-call one; oracle spot checks."""
    print("[gpu] decode_volume 512^3 b64", flush=True)
    data = fcs.encode_generated(fcs.Generator(2), 64)
    ds = O.Dataset(data)
    ns = np.array([ds.brick(i).nsym for i in range(ds.num_bricks)])
    n_long = int(np.sum(2 * ns >= ns.max()))
    assert 0 < n_long < ds.num_bricks
    st, got = _full(ctx, data, 0, True)
    _, want = _full(ctx, data, 0, False)
    assert np.all(st == 0)
    assert np.array_equal(got, want)
    for i in (int(np.argmax(ns)), int(np.argmin(ns)), 100, 511):
        lab, s = ds.decode(i, ds.L)
        assert s == 0
        assert np.array_equal(got[i, O.pyramid_offset(ds.L):].astype(np.uint64), lab)


@pytest.mark.parametrize("name", ["rand_b8", "cfg1_b32"])
def test_decode_volume_direct_pool(ctx, manifest, name):
    """A direct-mapped pool (brick i at slot i, slots of the finest level's
    size): single-level decodes at the finest LOD and one coarser, as the
    two-call path writes them."""
    print(f"[gpu] decode_volume direct pool {name}", flush=True)
    b = manifest[name]["brick_size"]
    data = golden_bytes(name)
    for lod_drop in (0, 1):
        out = []
        for pipelined in (True, False):
            vol = ctx.load(data, b)
            try:
                vol.cache_reserve_direct()
                L, nb = vol.info["levels"], vol.info["num_bricks"]
                t = L - lod_drop
                if pipelined:
                    assert np.all(vol.decode_volume(t, False) == 0)
                else:
                    vol.decode_rans()
                    vol.decode_all(t, False)
                stride, n = 1 << (3 * L), 1 << (3 * t)
                out.append(np.stack([vol.read_cache(i * stride, n) for i in range(nb)]))
            finally:
                vol.close()
        assert np.array_equal(out[0], out[1]), (name, lod_drop)
        ds = O.Dataset(data, b)
        for i in (0, ds.num_bricks - 1):
            assert np.array_equal(out[0][i].astype(np.uint64), ds.decode(i, ds.L - lod_drop)[0])


def test_decode_volume_repeated_calls_and_residency_changes(ctx):
    """Repeated calls, and calls after a residency change in between --
    requests at another LOD, another LOD for the whole volume, a
    direct-mapped pool, a device-built frame that re-decodes bricks at their
    frame LOD, a full decode through decode_all, a pyramid-layout call: every
    call leaves every brick's finest labels equal to the oracle's and every
    brick resident at that LOD. (A call launches the requests the previous
    one staged when nothing restaged them and the LOD, layout and slot size
    are the same; the calls in between restage them or change the key.)"""
    print("[gpu] decode_volume repeated calls", flush=True)
    data = golden_bytes("cfg1_b32")
    ds = O.Dataset(data, 32)
    vol = ctx.load(data, 32)
    try:
        L, nb = vol.info["levels"], vol.info["num_bricks"]
        per = 1 << (3 * L)
        want = np.stack([ds.decode(i, L)[0] for i in range(nb)])

        def check(stride):
            assert np.all(vol.decode_volume(L, False) == 0)
            got = np.stack([vol.read_cache(i * stride, per) for i in range(nb)]).astype(np.uint64)
            assert np.array_equal(got, want)
            bricks, lods = vol.cache_resident()
            assert len(bricks) == nb and np.all(lods == L)

        for _ in range(3):
            check(per)
        vol.cache_clear()  # (a full decode leaves the pool exactly full)
        vol.request_bricks(np.arange(8, dtype=np.uint32), np.full(8, 2, np.uint8))
        check(per)
        assert np.all(vol.decode_volume(2, False) == 0)
        check(per)
        vol.cache_reserve_direct()
        check(per)
        check(per)
        cam = vol.orbit_camera(0.7, 0.3, 3.0 * vol.info["grid_size"], math.pi / 3)
        vol.render(cam, 160, 120, want=())  # device-built frame: bricks re-decoded at their frame LOD
        check(per)
        vol.decode_all(L, True)  # restages the context's requests
        check(per)
        assert np.all(vol.decode_volume(L, True) == 0)  # pyramid layout: other slots
        pyr = O.pyramid_offset(L + 1)
        got = np.stack([vol.read_cache(i * pyr + O.pyramid_offset(L), per) for i in range(nb)]).astype(np.uint64)
        assert np.array_equal(got, want)
        check(per)
        check(per)
    finally:
        vol.close()
