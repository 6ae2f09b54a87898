"""The CPU .csdb producer (fcs_producer.cpp) against the reference encoder:
committed golden bytes (written by the reference encoder) and, where the
reference library is built, random tie-heavy volumes encoded by both."""
import numpy as np
import pytest

import fcs
from conftest import GOLDEN, brick_from_volume, golden_bytes, golden_volume, sha1
from oracle import oracle as O

NAMES = ["cfg1_b32", "rand_b2", "rand_b4", "rand_b8", "const_mix_b8", "all_zero_b8", "checker_b16",
         "lookback15_b8", "lookback16_b8", "u64_b8"]


@pytest.mark.parametrize("name", NAMES)
def test_producer_bytes_equal_reference_encoder(name, manifest):
    vol = golden_volume(name, manifest)
    b = manifest[name]["brick_size"]
    data = fcs.encode_volume(vol.astype(np.uint64), b, header_brick_size_64=True)
    assert data == golden_bytes(name)


def test_producer_writes_true_brick_size(manifest):
    vol = golden_volume("rand_b8", manifest)
    a = fcs.encode_volume(vol.astype(np.uint64), 8)
    r = golden_bytes("rand_b8")
    assert a[12:16] == (8).to_bytes(4, "little") and r[12:16] == (64).to_bytes(4, "little")
    assert a[:12] == r[:12] and a[16:] == r[16:]


def test_u32_and_u64_inputs_encode_identically(manifest):
    vol = golden_volume("cfg1_b32", manifest)
    assert fcs.encode_volume(vol.astype(np.uint32), 32) == fcs.encode_volume(vol.astype(np.uint64), 32)


def test_generated_stream_equals_materialised(manifest):
    g = fcs.Generator(1)
    assert fcs.encode_generated(g, 32, header_brick_size_64=True) == golden_bytes("cfg1_b32")


@pytest.mark.parametrize("name", ["cfg1_b32", "checker_b16", "rand_b4", "const_mix_b8"])
def test_pyramid_matches_reference_pyramid(name, manifest):
    vol = golden_volume(name, manifest)
    b = manifest[name]["brick_size"]
    L = b.bit_length() - 1
    exp = manifest[name]["bricks"]
    for bid, e in exp.items():
        pyr = fcs.brick_pyramid(brick_from_volume(vol, int(bid), b), b)
        for t in range(L + 1):
            assert sha1(pyr[O.pyramid_offset(t):O.pyramid_offset(t + 1)]) == e["lod"][t]


def test_generator_recipes():
    g1 = fcs.Generator(1)
    assert g1.size == 128
    assert abs(g1.background_fraction() - 0.3) < 0.08
    g2 = fcs.Generator(2, size=64)  # config-2 recipe at a small edge
    v = g2.volume()
    assert v.shape == (64, 64, 64)
    assert abs(g2.background_fraction() - 0.25) < 0.05
    # deterministic
    assert np.array_equal(v, fcs.Generator(2, size=64).volume())
    g3 = fcs.Generator(3, size=512)  # regions recipe, 2x2x2 regions of 256^3
    box = g3.box(0, 0, 0, 32, 32, 32)
    assert box.shape == (32, 32, 32)


@pytest.mark.skipif(not O.ref_available(), reason="reference library not built here")
def test_random_volumes_against_reference_encoder(tmp_path):
    rng = np.random.default_rng(7)
    for trial in range(60):
        b = int(rng.choice([2, 4, 8]))
        dims = [int(rng.integers(1, 3 * b)) for _ in range(3)]
        k = int(rng.integers(1, 5))
        v = rng.integers(0, k, dims[::-1]).astype(np.uint64)
        ref, _ = O.ref_encode(v, b, str(tmp_path / "r.csbd"))
        assert fcs.encode_volume(v, b, header_brick_size_64=True) == ref, trial
