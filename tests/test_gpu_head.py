"""The op decoder's head pass (levels up to 3 decoded one wave per request,
state handed over through the request's output) against the oracle and
against the same decode without it, bit-exact: pyramids, single-LOD batches
of mixed LODs, both label widths, 64^3 bricks and corrupt nibble streams."""
import numpy as np
import pytest

import fcs
from conftest import golden_bytes
from oracle import oracle as O

pytestmark = pytest.mark.gpu

CASES = ["cfg1_b32", "rand_b2", "rand_b4", "rand_b8", "const_mix_b8", "all_zero_b8", "checker_b16",
         "lookback15_b8", "lookback16_b8", "u64_b8", "microns8_b64"]


@pytest.fixture(scope="module")
def ctx_on():
    if fcs.device_count() == 0:
        pytest.skip("no HIP device")
    c = fcs.Context(0)
    c.set_ops_head(fcs.OPS_HEAD_ON)
    yield c
    c.close()


@pytest.fixture(scope="module")
def ctx_off():
    if fcs.device_count() == 0:
        pytest.skip("no HIP device")
    c = fcs.Context(0)
    c.set_ops_head(fcs.OPS_HEAD_OFF)
    yield c
    c.close()


def _pyramids(ctx, data, b):
    vol = ctx.load(data, b)
    vol.decode_rans()
    vol.decode_all(pyramid=True)
    L, nb = vol.info["levels"], vol.info["num_bricks"]
    per = O.pyramid_offset(L + 1)
    cache = vol.read_cache(0, per * nb).reshape(nb, per).astype(np.uint64)
    vol.close()
    return cache


@pytest.mark.parametrize("name", CASES)
def test_head_pyramids_match_oracle_and_plain_decode(ctx_on, ctx_off, manifest, name):
    print(f"[gpu] head pass pyramids {name}", flush=True)
    b = manifest[name]["brick_size"]
    data = golden_bytes(name)
    on = _pyramids(ctx_on, data, b)
    off = _pyramids(ctx_off, data, b)
    assert np.array_equal(on, off)
    ds = O.Dataset(data, b)
    L = ds.L
    for i in range(ds.num_bricks):
        for t in range(L + 1):
            want, s = ds.decode(i, t)
            assert s == 0
            assert np.array_equal(on[i, O.pyramid_offset(t):O.pyramid_offset(t + 1)], want), (name, i, t)


@pytest.mark.parametrize("name", ["cfg1_b32", "u64_b8", "microns8_b64"])
def test_head_mixed_lod_batch(ctx_on, manifest, name):
    print(f"[gpu] head pass mixed LODs {name}", flush=True)
    b = manifest[name]["brick_size"]
    data = golden_bytes(name)
    ds = O.Dataset(data, b)
    vol = ctx_on.load(data, b)
    nb, L = vol.info["num_bricks"], vol.info["levels"]
    rng = np.random.default_rng(11)
    bricks = rng.permutation(nb).astype(np.uint32)
    lods = rng.integers(0, L + 1, nb).astype(np.uint8)
    off, st = vol.request_bricks(bricks, lods)
    assert np.all(st == 0)
    for k in range(nb):
        i, t = int(bricks[k]), int(lods[k])
        got = vol.read_cache(int(off[k]), 1 << (3 * t)).astype(np.uint64)
        assert np.array_equal(got, ds.decode(i, t)[0]), (name, i, t)
    vol.close()


def test_head_corrupt_streams_same_as_plain_decode(ctx_on, ctx_off):
    """Bricks whose rANS streams were damaged decode to garbage nibbles; both
    paths must report the same statuses and write the same labels."""
    print("[gpu] head pass corrupt streams", flush=True)
    data = bytearray(golden_bytes("rand_b8"))
    ds = O.Dataset(bytes(data), 8)
    for i in range(0, ds.num_bricks, 3):
        br = ds.brick(i)
        if br.E > 8:
            data[br.str_off + br.E // 3] ^= 0x5A
    res = []
    for c in (ctx_on, ctx_off):
        vol = c.load(bytes(data), 8)
        vol.decode_rans(raise_on_error=False)
        L, nb = vol.info["levels"], vol.info["num_bricks"]
        vol.cache_reserve(nb << (3 * L))
        off, st = vol.request_bricks(np.arange(nb, dtype=np.uint32), np.full(nb, L, np.uint8),
                                     raise_on_error=False)
        n = 1 << (3 * L)
        res.append((st.copy(), np.stack([vol.read_cache(int(o), n) for o in off])))
        vol.close()
    assert np.any(res[0][0] != 0)
    assert np.array_equal(res[0][0], res[1][0])
    assert np.array_equal(res[0][1], res[1][1])


def test_head_auto_on_large_batch_matches_plain(ctx_off):
    """A 512^3 volume of 32^3 bricks (4096 requests: the automatic mode uses
    the head pass) decodes to the same finest level as without it."""
    print("[gpu] head pass auto, 512^3 b32", flush=True)
    g = fcs.Generator(3, size=512)
    data = fcs.encode_generated(g, 32)
    auto = fcs.Context(0)
    try:
        outs = []
        for c in (auto, ctx_off):
            vol = c.load(data)
            vol.decode_rans()
            nb, L = vol.info["num_bricks"], vol.info["levels"]
            vol.cache_reserve(nb * (1 << (3 * L)))
            vol.decode_all(lod=L, pyramid=False)
            outs.append(vol.read_cache(0, nb << (3 * L)).copy())
            vol.close()
        assert np.array_equal(outs[0], outs[1])
        # spot-check against the oracle
        ds = O.Dataset(data)
        n = 1 << (3 * ds.L)
        for i in (0, 1, 77, 1000, 4095):
            assert np.array_equal(outs[0][i * n:(i + 1) * n].astype(np.uint64), ds.decode(i, ds.L)[0])
    finally:
        auto.close()
