"""Pins the CPU oracle (oracle/oracle.c) against what the reference's own code
produced (tests/golden, written by tests/golden/make_golden.py), against the
source volumes, and against closed forms; then shows that plausible mistakes
planted in the oracle fail these checks."""
import os
import subprocess

import numpy as np
import pytest

from conftest import GOLDEN, ROOT, brick_from_volume, golden_bytes, golden_volume, sha1
from oracle import oracle as O

CASES = ["cfg1_b32", "rand_b2", "rand_b4", "rand_b8", "const_mix_b8", "all_zero_b8", "checker_b16",
         "lookback15_b8", "lookback16_b8", "u64_b8", "microns8_b64"]


def _ds(name, manifest):
    return O.Dataset(golden_bytes(name), manifest[name]["brick_size"])


@pytest.mark.parametrize("name", CASES)
def test_rans_matches_reference_decoder(name, manifest):
    ds = _ds(name, manifest)
    exp = manifest[name]["bricks"]
    assert ds.num_bricks == len(exp)
    for i in range(ds.num_bricks):
        nib, st = ds.rans(i)
        assert st == 0, (name, i)
        assert sha1(nib) == exp[str(i)]["rans"], (name, i)


@pytest.mark.parametrize("name", [c for c in CASES if c != "microns8_b64"])
def test_every_lod_matches_encoder_pyramid(name, manifest):
    ds = _ds(name, manifest)
    exp = manifest[name]["bricks"]
    for i in range(ds.num_bricks):
        for t in range(ds.L + 1):
            out, st = ds.decode(i, t)
            assert st == 0
            assert sha1(out) == exp[str(i)]["lod"][t], (name, i, t)


@pytest.mark.parametrize("name", CASES)
def test_finest_matches_reference_decodebrick(name, manifest):
    ds = _ds(name, manifest)
    exp = manifest[name]["bricks"]
    checked = 0
    for i in range(ds.num_bricks):
        if exp[str(i)]["ref_finest"] is None:
            continue  # the reference decoder throws on constant bricks
        out, st = ds.decode(i, ds.L)
        assert st == 0
        assert sha1(out) == exp[str(i)]["ref_finest"], (name, i)
        checked += 1
    if name not in ("all_zero_b8",):
        assert checked > 0


@pytest.mark.parametrize("name", [c for c in CASES if c != "microns8_b64"])
def test_finest_equals_source_volume(name, manifest):
    vol = golden_volume(name, manifest)
    assert vol is not None
    ds = _ds(name, manifest)
    b = ds.brick_size
    for i in range(ds.num_bricks):
        out, _ = ds.decode(i, ds.L)
        assert np.array_equal(out, brick_from_volume(vol, i, b)), (name, i)


@pytest.mark.parametrize("name", ["cfg1_b32", "const_mix_b8", "u64_b8", "microns8_b64"])
def test_lod0_is_first_palette_entry(name, manifest):
    ds = _ds(name, manifest)
    for i in range(ds.num_bricks):
        out, st = ds.decode(i, 0)
        assert out[0] == ds.palette(i)[0]


def test_constant_bricks_fill_every_lod(manifest):
    ds = _ds("const_mix_b8", manifest)
    n_const = 0
    for i in range(ds.num_bricks):
        nib, _ = ds.rans(i)
        if ds.brick(i).nsym == 1:
            assert nib[0] >> 4 & 1  # root stop bit set
            n_const += 1
            for t in range(ds.L + 1):
                out, st = ds.decode(i, t)
                assert st == 0 and np.all(out == ds.palette(i)[0])
    assert n_const >= 6


def test_lookback_fixtures_exercise_delta15_and_readd(manifest):
    def deltas(name):
        ds = _ds(name, manifest)
        nib, _ = ds.rans(0)
        nibs = np.stack([nib >> 4, nib & 15], 1).reshape(-1)[: ds.brick(0).nsym]
        out, j = [], 1
        while j < len(nibs):
            if nibs[j] >> 1 == 6:
                out.append(int(nibs[j + 1]))
                j += 1
            j += 1
        return out, ds.palette(0)

    d15, p15 = deltas("lookback15_b8")
    d16, p16 = deltas("lookback16_b8")
    assert 15 in d15
    assert 15 not in d16 and len(p16) != len(np.unique(p16))  # re-added duplicate


def test_rans_invariant_flags_corruption(manifest):
    data = bytearray(golden_bytes("rand_b8"))
    ds = O.Dataset(bytes(data), 8)
    b = ds.brick(3)
    data[b.str_off + b.E // 2] ^= 0x5A
    bad = O.Dataset(bytes(data), 8)
    _, st = bad.rans(3)
    assert st & 16


def test_parse_rejects_bad_magic():
    with pytest.raises(ValueError, match="Invalid file format"):
        O.Dataset(b"XXXX" + b"\0" * 200)


def test_ortho_render_matches_column_scan(manifest):
    """Orthographic +z frame over config 1 at the finest LOD: every pixel's
    first hit equals a brute-force scan of its voxel column."""
    vol = golden_volume("cfg1_b32", manifest)
    ds = _ds("cfg1_b32", manifest)
    B, b, G = 4, 32, 128
    labels = [ds.decode(i, 5)[0] for i in range(ds.num_bricks)]
    empty = [int(np.all(ds.palette(i) == 0)) for i in range(ds.num_bricks)]
    p0 = [ds.palette(i)[0] for i in range(ds.num_bricks)]
    cam = {"position": [64.0, 64.0, -1.0], "forward": [0, 0, 1.0], "right": [1.0, 0, 0], "up": [0, -1.0, 0],
           "tan_half_fov": 0.0, "ortho_half_height": 64.0}
    W = H = 256
    lab, vox, tt, vis = O.render(B, b, empty, [5] * 64, labels, p0, cam, W, H)
    px = (np.arange(W) + 0.5) * G / W
    for py in range(0, H, 7):
        y = int((py + 0.5) * G / H)
        for pxi in range(0, W, 5):
            x = int(px[pxi])
            col = vol[:, y, x]
            nz = np.nonzero(col)[0]
            k = py * W + pxi
            if len(nz) == 0:
                assert lab[k] == 0 and vox[k] == np.uint64(2 ** 64 - 1)
            else:
                z = nz[0]
                assert lab[k] == col[z]
                assert vox[k] == x + G * (y + G * z)


def test_ortho_render_coarse_lod_matches_cell_scan(manifest):
    """Same at LOD 3 (stride 4): the hit is the first nonzero LOD-3 cell of the
    column, reported at the cell's min corner."""
    ds = _ds("cfg1_b32", manifest)
    B, b, G, t = 4, 32, 128, 3
    s = b >> t
    labels = [ds.decode(i, t)[0] for i in range(ds.num_bricks)]
    empty = [int(np.all(ds.palette(i) == 0)) for i in range(ds.num_bricks)]
    p0 = [ds.palette(i)[0] for i in range(ds.num_bricks)]
    cam = {"position": [64.0, 64.0, -1.0], "forward": [0, 0, 1.0], "right": [1.0, 0, 0], "up": [0, -1.0, 0],
           "tan_half_fov": 0.0, "ortho_half_height": 64.0}
    W = H = 128
    lab, vox, _, _ = O.render(B, b, empty, [t] * 64, labels, p0, cam, W, H)

    def mort(x, y, z):
        out = 0
        for i in range(10):
            out |= ((x >> i) & 1) << (3 * i) | ((y >> i) & 1) << (3 * i + 1) | ((z >> i) & 1) << (3 * i + 2)
        return out

    for py in range(0, H, 9):
        for pxi in range(0, W, 11):
            x, y = pxi, py
            exp_lab, exp_vox = 0, 2 ** 64 - 1
            for z in range(0, G, s):
                bi = mort(x // b, y // b, z // b)
                if empty[bi]:
                    continue
                v = labels[bi][mort((x % b) // s, (y % b) // s, (z % b) // s)]
                if v != 0:
                    exp_lab = v
                    qx, qy = (x // s) * s, (y // s) * s
                    exp_vox = qx + G * (qy + G * z)
                    break
            k = py * W + pxi
            assert lab[k] == exp_lab and int(vox[k]) == exp_vox, (pxi, py)


def test_lod_rule_thresholds():
    # focal = (sw/2)/tan(fov/2); LOD t valid while focal * (b / 2^t) / d >= 1
    fov = 2 * np.arctan(0.5)  # tan(fov/2) = 0.5 -> focal = sw
    sw, b = 100.0, 32.0
    # d = focal * b / 2^t exactly at the threshold keeps LOD t
    for t in range(6):
        d = sw * b / 2 ** t
        assert O.brick_lod((d, 0, 0), b, (0, 0, 0), fov, sw, 5) == t
        assert O.brick_lod((d * 1.0001, 0, 0), b, (0, 0, 0), fov, sw, 5) == max(t - 1, 0)
    assert O.brick_lod((0.01, 0, 0), b, (0, 0, 0), fov, sw, 5) == 5  # clamped distance 0.1


MUTANTS = {1: "neighbour after in Morton order read at the child level",
           2: "Pd delta off by one",
           3: "root stop bit ignored",
           4: "LOD quantisation of the voxel DDA dropped",
           5: "Pa does not advance the palette cursor"}


@pytest.mark.parametrize("mutant", sorted(MUTANTS))
def test_planted_oracle_mistakes_are_caught(mutant, manifest, tmp_path):
    so = tmp_path / f"liboracle_m{mutant}.so"
    subprocess.run(["gcc", "-O2", "-fPIC", "-std=c11", "-ffp-contract=off", "-shared", f"-DOR_MUTANT={mutant}",
                    "-o", str(so), os.path.join(ROOT, "oracle", "oracle.c"), "-lm", "-lpthread"], check=True)
    import ctypes

    saved = O._lib
    try:
        O._lib = None
        real = O.ORACLE_SO
        O.ORACLE_SO = str(so)
        caught = False
        try:
            for name in ("cfg1_b32", "const_mix_b8", "lookback15_b8", "rand_b4"):
                try:
                    test_every_lod_matches_encoder_pyramid(name, manifest)
                except AssertionError:
                    caught = True
                    break
            if not caught:
                try:
                    test_ortho_render_coarse_lod_matches_cell_scan(manifest)
                except AssertionError:
                    caught = True
        finally:
            O.ORACLE_SO = real
        assert caught, MUTANTS[mutant]
    finally:
        O._lib = saved
