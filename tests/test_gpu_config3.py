"""The batching workload at full size: the 2048^3 regions volume (seed 2:
256^3 regions of background / ~50 cells / ~4000 noisy cells), 32^3 bricks
(262,144 bricks), encoded by the CPU producer.

* fcs_decode_volume (the rANS stage and the finest-LOD decode of every brick,
  both in one call) is compared with the generator's
  source bricks, every label of every brick (closed form: the finest LOD of a
  decoded brick is the source volume; decode.wgsl:344-508 and
  brickDecoding.hpp:174-242 reproduce the encoded brick exactly).
* One batched launch of 100k shuffled (brick, LOD uniform in 0..5) requests
  is compared with the oracle's decode of the same requests, request by
  request (decode.wgsl's level truncation for the coarse LODs).
"""
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

SIZE, B = 2048, 32
CHUNK = 8192  # bricks per host comparison (1 GiB of labels)


def say(msg):
    print(msg, file=sys.stderr, flush=True)


@pytest.fixture(scope="module")
def cfg3(gpu_ctx):
    import fcs

    gen = fcs.Generator(3)
    assert gen.size == SIZE
    say("[config3] encoding the 2048^3 regions volume")
    data = fcs.encode_generated(gen, B)
    vol = gpu_ctx.load(data)
    yield gen, data, vol
    vol.close()


def test_full_decode_equals_source_volume(cfg3):
    gen, _, vol = cfg3
    info = vol.info
    nb, L = info["num_bricks"], info["levels"]
    assert nb == (SIZE // B) ** 3 and L == 5
    st = vol.decode_volume(lod=L, pyramid=False)
    assert not np.any(st), f"{int(np.count_nonzero(st))} bricks reported a status"
    # every brick is resident at its finest LOD: the offsets come back without a decode
    off, st2 = vol.request_bricks(np.arange(nb, dtype=np.uint32), np.full(nb, L, np.uint8))
    assert not np.any(st2)
    per = 1 << (3 * L)
    assert np.array_equal(off, np.arange(nb, dtype=np.uint64) * np.uint64(per)), "bricks are not slot-ordered"
    bad = []
    for first in range(0, nb, CHUNK):
        n = min(CHUNK, nb - first)
        dev = vol.read_cache(first * per, n * per).reshape(n, per)
        src = gen.bricks(B, first, n)
        if not np.array_equal(dev, src):
            rows = np.nonzero(np.any(dev != src, axis=1))[0]
            bad.extend(int(first + r) for r in rows[:8])
        say(f"[config3] bricks {first + n} of {nb} compared with the source, {len(bad)} differing so far")
    assert not bad, f"bricks differing from the source volume: {bad[:16]}"


def test_batched_requests_equal_oracle(cfg3):
    from oracle import oracle as O

    _, data, vol = cfg3
    info = vol.info
    nb, L = info["num_bricks"], info["levels"]
    rng = np.random.default_rng(2)
    nreq = 100000
    bricks = rng.choice(nb, nreq, replace=False).astype(np.uint32)
    lods = rng.integers(0, L + 1, nreq).astype(np.uint8)
    vol.cache_clear()
    off, st = vol.request_bricks(bricks, lods)
    assert not np.any(st)
    sizes = np.left_shift(np.uint64(1), np.uint64(3) * lods.astype(np.uint64))
    end = int((off + sizes).max())
    say(f"[config3] {nreq} batched requests decoded, reading {end} labels back")
    dev = vol.read_cache(0, end)
    ds = O.Dataset(data)
    ref, roff, rst = ds.decode_requests(bricks, lods, np.uint32)
    assert not np.any(rst)
    say("[config3] oracle decoded the same requests")
    bad = []
    for k in range(nreq):
        n = int(sizes[k])
        a = dev[int(off[k]):int(off[k]) + n]
        r = ref[int(roff[k]):int(roff[k]) + n]
        if not np.array_equal(a, r):
            bad.append((int(bricks[k]), int(lods[k])))
            if len(bad) >= 16:
                break
    assert not bad, f"requests differing from the oracle: {bad}"
