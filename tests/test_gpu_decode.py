"""Device decoder parity through the C ABI: rANS stage and op decoder against
the oracle and the reference-produced golden expectations, bit-exact."""
import numpy as np
import pytest

import fcs
from conftest import brick_from_volume, golden_bytes, sha1
from oracle import oracle as O

pytestmark = pytest.mark.gpu

CASES = ["cfg1_b32", "rand_b2", "rand_b4", "rand_b8", "const_mix_b8", "all_zero_b8", "checker_b16",
         "lookback15_b8", "lookback16_b8", "u64_b8", "microns8_b64"]


def _load(ctx, name, manifest):
    b = manifest[name]["brick_size"]
    return ctx.load(golden_bytes(name), b), O.Dataset(golden_bytes(name), b)


@pytest.mark.parametrize("name", CASES)
def test_pyramid_decode_matches_oracle_and_reference(gpu_ctx, manifest, name):
    print(f"[gpu] pyramid decode {name}", flush=True)
    vol, ds = _load(gpu_ctx, name, manifest)
    st = vol.decode_rans()
    assert np.all(st == 0)
    vol.decode_all(pyramid=True)
    L = vol.info["levels"]
    per = O.pyramid_offset(L + 1)
    cache = vol.read_cache(0, per * vol.info["num_bricks"]).reshape(-1, per)
    exp = manifest[name]["bricks"]
    for i in range(vol.info["num_bricks"]):
        for t in range(L + 1):
            got = cache[i, O.pyramid_offset(t):O.pyramid_offset(t + 1)].astype(np.uint64)
            want, s = ds.decode(i, t)
            assert s == 0
            assert np.array_equal(got, want), (name, i, t)
            if "lod" in exp[str(i)]:
                assert sha1(got) == exp[str(i)]["lod"][t]
        if exp[str(i)]["ref_finest"] is not None:
            assert sha1(cache[i, O.pyramid_offset(L):].astype(np.uint64)) == exp[str(i)]["ref_finest"]
    vol.close()


def test_label_width_follows_palette(gpu_ctx, manifest):
    v64, _ = _load(gpu_ctx, "u64_b8", manifest)
    v32, _ = _load(gpu_ctx, "rand_b8", manifest)
    assert v64.info["label_bits"] == 64 and v32.info["label_bits"] == 32


def test_single_lod_requests_and_cache_hits(gpu_ctx, manifest):
    print("[gpu] request_bricks cfg1", flush=True)
    vol, ds = _load(gpu_ctx, "cfg1_b32", manifest)
    rng = np.random.default_rng(3)
    bricks = rng.integers(0, 64, 200).astype(np.uint32)
    lods = rng.integers(0, 6, 200).astype(np.uint8)
    off, st = vol.request_bricks(bricks, lods)
    assert np.all(st == 0)
    for k in range(200):
        b, t = int(bricks[k]), int(lods[k])
        # a later request for the same brick at another LOD evicts this one
        later = [j for j in range(k + 1, 200) if bricks[j] == b and lods[j] != t]
        if later:
            continue
        got = vol.read_cache(int(off[k]), 1 << (3 * t)).astype(np.uint64)
        assert np.array_equal(got, ds.decode(b, t)[0]), (b, t)
    # resident bricks are not decoded again: same offsets
    off2, st2 = vol.request_bricks(bricks[-10:], lods[-10:])
    assert np.all(st2 == 0) and np.array_equal(off2, off[-10:])
    # read_brick == oracle at every LOD
    for t in range(6):
        assert np.array_equal(vol.read_brick(13, t).astype(np.uint64), ds.decode(13, t)[0])


@pytest.mark.parametrize("t", [0, 2, 5])
def test_decode_all_single_lod(gpu_ctx, manifest, t):
    vol, ds = _load(gpu_ctx, "cfg1_b32", manifest)
    vol.decode_all(lod=t, pyramid=False)
    n = 1 << (3 * t)
    cache = vol.read_cache(0, n * 64).reshape(64, n)
    for i in range(64):
        assert np.array_equal(cache[i].astype(np.uint64), ds.decode(i, t)[0])


def test_corrupt_stream_is_flagged_per_brick(gpu_ctx, manifest):
    data = bytearray(golden_bytes("rand_b8"))
    ds = O.Dataset(bytes(data), 8)
    b = ds.brick(5)
    data[b.str_off + b.E // 2] ^= 0xA5
    vol = gpu_ctx.load(bytes(data), 8)
    st = vol.decode_rans(raise_on_error=False)
    assert st[5] != 0
    assert np.all(np.delete(st, 5) == 0)


def test_cache_full_is_reported(gpu_ctx, manifest):
    vol, _ = _load(gpu_ctx, "cfg1_b32", manifest)
    vol.cache_reserve(2 * 32768)
    off, st = vol.request_bricks(np.arange(4, dtype=np.uint32), np.full(4, 5, np.uint8), raise_on_error=False)
    assert list(st[:2]) == [0, 0]
    assert np.all(st[2:] & fcs.ST_CACHE_FULL)


def test_bad_magic_rejected(gpu_ctx):
    with pytest.raises(fcs.FcsError, match="Invalid file format"):
        gpu_ctx.load(b"ABCD" + bytes(300))


@pytest.fixture(scope="module")
def cfg2():
    print("[gpu] generating config 2 (512^3, 64^3 bricks)", flush=True)
    g = fcs.Generator(2)
    vol = g.volume()
    data = fcs.encode_volume(vol, 64)
    return vol, data


def test_config2_full_decode_bit_exact(gpu_ctx, cfg2):
    src, data = cfg2
    print("[gpu] config 2 full pyramid decode", flush=True)
    vol = gpu_ctx.load(data)
    assert np.all(vol.decode_rans() == 0)
    vol.decode_all(pyramid=True)
    per = O.pyramid_offset(7)
    cache = vol.read_cache(0, per * 512).reshape(512, per)
    ds = O.Dataset(data)
    ref, s = ds.decode_all_pyramid()
    assert s == 0
    assert np.array_equal(cache.astype(np.uint64), ref)
    for i in (0, 77, 300, 511):
        assert np.array_equal(cache[i, O.pyramid_offset(6):].astype(np.uint64), brick_from_volume(src, i, 64))


@pytest.mark.parametrize("mapping", [1, 2, 3, 4, 5, 6])
@pytest.mark.parametrize("name", ["cfg1_b32", "microns8_b64", "checker_b16", "const_mix_b8", "rand_b2", "lookback16_b8"])
def test_both_rans_mappings_bit_exact(gpu_ctx, manifest, name, mapping):
    """Lane-per-brick (1) and wave-per-brick (2: scalar table, 3: ballot, 4: LDS table, 5/6: scalar byte
    window with table / ballot) rANS kernels, forced."""
    print(f"[gpu] rANS mapping {mapping} on {name}", flush=True)
    gpu_ctx.set_rans_mapping(mapping)
    try:
        vol, ds = _load(gpu_ctx, name, manifest)
        assert np.all(vol.decode_rans() == 0)
        assert gpu_ctx.timing()["rans_mapping"] == mapping
        vol.decode_all(pyramid=True)
        L = vol.info["levels"]
        per = O.pyramid_offset(L + 1)
        cache = vol.read_cache(0, per * vol.info["num_bricks"]).reshape(-1, per).astype(np.uint64)
        ref, s = ds.decode_all_pyramid()
        assert s == 0 and np.array_equal(cache, ref)
    finally:
        gpu_ctx.set_rans_mapping(0)


def test_config2_both_mappings(gpu_ctx, cfg2):
    _, data = cfg2
    outs = []
    for mapping in (1, 2, 3, 4, 5, 6):
        gpu_ctx.set_rans_mapping(mapping)
        vol = gpu_ctx.load(data)
        assert np.all(vol.decode_rans() == 0)
        vol.decode_all(lod=6, pyramid=False)
        outs.append(vol.read_cache(0, 512 * 262144))
        vol.close()
    gpu_ctx.set_rans_mapping(0)
    assert all(np.array_equal(outs[0], o) for o in outs[1:])


@pytest.mark.parametrize("split", [1, 2])
@pytest.mark.parametrize("name", ["cfg1_b32", "rand_b8", "checker_b16", "u64_b8", "lookback16_b8"])
def test_ops_lds_splits_bit_exact(gpu_ctx, manifest, name, split):
    """Both op decoder LDS splits for bricks up to 32^3 (1: three workgroups
    per CU, 2: four), forced, against the oracle's pyramid."""
    print(f"[gpu] ops split {split} on {name}", flush=True)
    gpu_ctx.set_ops_split(split)
    try:
        vol, ds = _load(gpu_ctx, name, manifest)
        assert np.all(vol.decode_rans() == 0)
        vol.decode_all(pyramid=True)
        assert gpu_ctx.timing()["ops_split"] == split
        L = vol.info["levels"]
        per = O.pyramid_offset(L + 1)
        cache = vol.read_cache(0, per * vol.info["num_bricks"]).reshape(-1, per).astype(np.uint64)
        ref, s = ds.decode_all_pyramid()
        assert s == 0 and np.array_equal(cache, ref)
        vol.close()
    finally:
        gpu_ctx.set_ops_split(0)


def test_ops_split_autotune_512cube_b32(gpu_ctx, cfg2):
    """512^3 re-encoded with 32^3 bricks (4,096 requests): the first large
    batch times both splits and keeps the faster; the labels equal those of
    either forced split and the source volume."""
    src, _ = cfg2
    data = fcs.encode_volume(src, 32)
    outs = {}
    for split in (0, 1, 2):
        print(f"[gpu] 512^3 b32 full decode, ops split {split}", flush=True)
        gpu_ctx.set_ops_split(split)
        vol = gpu_ctx.load(data)
        assert np.all(vol.decode_rans() == 0)
        vol.decode_all(lod=5, pyramid=False)
        t = gpu_ctx.timing()
        if split == 0:
            assert t["tune_ops_3wg_ms"] > 0 and t["tune_ops_4wg_ms"] > 0
            assert t["ops_split"] in (1, 2)
        else:
            assert t["ops_split"] == split
        outs[split] = vol.read_cache(0, 4096 * 32768)
        vol.close()
    gpu_ctx.set_ops_split(0)
    assert np.array_equal(outs[0], outs[1]) and np.array_equal(outs[0], outs[2])
    cache = outs[0].reshape(4096, 32768)
    for i in (0, 1000, 4095):
        assert np.array_equal(cache[i].astype(np.uint64), brick_from_volume(src, i, 32))


@pytest.mark.parametrize("mapping", [1, 2, 5])
def test_rans_collapsed_state_reads_past_two_bytes(gpu_ctx, mapping):
    """A stream whose final state bytes are zeroed starts from state 0: the
    first symbols renormalise with more than the two bytes a valid stream can
    need per symbol (the lane decoder redoes such a brick by the sequential
    rule). Nibbles and statuses still equal the oracle's sequential decode."""
    print(f"[gpu] collapsed-state streams, mapping {mapping}", flush=True)
    data = bytearray(golden_bytes("cfg1_b32"))
    ds0 = O.Dataset(bytes(data), 32)
    nb = ds0.num_bricks
    hit = [i for i in range(nb) if ds0.brick(i).E > 16][::7]
    assert hit
    for i in hit:
        br = ds0.brick(i)
        data[br.str_off + br.E - 4:br.str_off + br.E] = bytes(4)
    ds = O.Dataset(bytes(data), 32)
    gpu_ctx.set_rans_mapping(mapping)
    try:
        vol = gpu_ctx.load(bytes(data), 32)
        st = vol.decode_rans(raise_on_error=False)
        for i in range(nb):
            want, wst = ds.rans(i)
            got = vol.read_nibbles(i)
            assert np.array_equal(got, want[:got.size]) and got.size == (ds.brick(i).nsym + 1) // 2, i
            assert (st[i] != 0) == (wst != 0), (i, st[i], wst)
        assert all(st[i] != 0 for i in hit)
        vol.close()
    finally:
        gpu_ctx.set_rans_mapping(0)

@pytest.mark.parametrize("mapping", [1, 2, 5])
def test_rans_initial_state_above_2_31(gpu_ctx, mapping):
    """A stream whose last byte (the top byte of the initial state) is 0xFF
    starts from a state of at least 2^31, which no valid stream reaches; the
    first symbols then leave states that need no renormalisation or wrap the
    lane decoder's byte count, so the lane decodes the brick again by the
    sequential rule. Nibbles and statuses equal the oracle's decode."""
    print(f"[gpu] initial state >= 2^31, mapping {mapping}", flush=True)
    data = bytearray(golden_bytes("cfg1_b32"))
    ds0 = O.Dataset(bytes(data), 32)
    nb = ds0.num_bricks
    hit = [i for i in range(nb) if ds0.brick(i).E > 16][3::5]
    assert hit
    for i in hit:
        br = ds0.brick(i)
        data[br.str_off + br.E - 1] = 0xFF
    ds = O.Dataset(bytes(data), 32)
    gpu_ctx.set_rans_mapping(mapping)
    try:
        vol = gpu_ctx.load(bytes(data), 32)
        st = vol.decode_rans(raise_on_error=False)
        for i in range(nb):
            want, wst = ds.rans(i)
            got = vol.read_nibbles(i)
            assert np.array_equal(got, want[:got.size]) and got.size == (ds.brick(i).nsym + 1) // 2, i
            assert (st[i] != 0) == (wst != 0), (i, st[i], wst)
        assert all(st[i] != 0 for i in hit)
        vol.close()
    finally:
        gpu_ctx.set_rans_mapping(0)


@pytest.mark.parametrize("mapping", [1, 2, 3, 4, 5, 6])
def test_rans_nibbles_equal_oracle_with_corrupt_bricks(gpu_ctx, manifest, mapping):
    """Every rANS mapping's nibble streams (fcs_read_nibbles) equal the
    oracle's decodeRansStream output byte for byte, with per-brick statuses,
    also for bricks whose stream was corrupted (flipped bytes near the start,
    middle and end) -- the byte-exact end-of-stream paths."""
    print(f"[gpu] nibbles vs oracle, mapping {mapping}", flush=True)
    for name, b in (("cfg1_b32", 32), ("microns8_b64", 64)):
        data = bytearray(golden_bytes(name))
        ds0 = O.Dataset(bytes(data), b)
        nb = ds0.num_bricks
        bad = [1, nb // 2, nb - 1]
        for k, i in enumerate(bad):
            br = ds0.brick(i)
            if br.E > 8:
                data[br.str_off + [1, br.E // 2, br.E - 6][k]] ^= 0x5A
        ds = O.Dataset(bytes(data), b)
        gpu_ctx.set_rans_mapping(mapping)
        try:
            vol = gpu_ctx.load(bytes(data), b)
            st = vol.decode_rans(raise_on_error=False)
            assert gpu_ctx.timing()["rans_mapping"] == mapping
            for i in range(nb):
                want, wst = ds.rans(i)
                got = vol.read_nibbles(i)
                assert np.array_equal(got, want[:got.size]) and got.size == (ds.brick(i).nsym + 1) // 2, (name, i)
                assert (st[i] != 0) == (wst != 0), (name, i, st[i], wst)
            vol.close()
        finally:
            gpu_ctx.set_rans_mapping(0)

