"""Device DDA marcher parity through the C ABI: per pixel first-hit label,
voxel index and hit t against the oracle's FP32 DDA, bit-exact (mismatching
pixels are counted and reported; the bar is zero)."""
import math

import numpy as np
import pytest

import fcs
from conftest import golden_bytes
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def oracle_frame(ds, pyr, lods, cam, w, h):
    B = round(ds.num_bricks ** (1 / 3))
    labels, empty, p0 = [], [], []
    for i in range(ds.num_bricks):
        t = int(lods[i])
        labels.append(pyr[i][O.pyramid_offset(t):O.pyramid_offset(t + 1)])
        pal = ds.palette(i)
        empty.append(int(np.all(pal == 0)))
        p0.append(pal[0])
    return O.render(B, ds.brick_size, empty, lods, labels, p0, cam, w, h)


def compare(frame, ref, tag):
    lab, vox, tt, vis = ref
    bad = (frame.label.astype(np.uint64) != lab) | (frame.voxel != vox) | (frame.t_hit.view(np.uint32) != tt.view(np.uint32))
    nbad = int(bad.sum())
    print(f"[gpu] {tag}: {frame.label.size} px, {nbad} mismatching, hits {int((lab != 0).sum())}", flush=True)
    assert nbad == 0
    assert np.array_equal(frame.brick_visited, vis)


@pytest.fixture(scope="module")
def cfg1(gpu_ctx):
    data = golden_bytes("cfg1_b32")
    vol = gpu_ctx.load(data, 32)
    vol.decode_all(pyramid=True)
    ds = O.Dataset(data, 32)
    pyr, _ = ds.decode_all_pyramid()
    return vol, ds, pyr


@pytest.mark.parametrize("lod", [5, 3, 1])
def test_ortho_frame_config1(cfg1, lod):
    vol, ds, pyr = cfg1
    cam = fcs.Camera.ortho_plus_z(128)
    lods = np.full(64, lod, np.uint8)
    f = vol.render(cam, 512, 512, brick_lod=lods)
    assert f.bricks_decoded == 0  # pyramid resident
    compare(f, oracle_frame(ds, pyr, lods, cam.as_dict(), 512, 512), f"ortho cfg1 lod {lod}")


def test_perspective_mixed_lods_config1(cfg1):
    vol, ds, pyr = cfg1
    rng = np.random.default_rng(11)
    lods = rng.integers(0, 6, 64).astype(np.uint8)
    for k in range(3):
        cam = vol.orbit_camera(0.3 + k, 0.2 * (k - 1), 1.5 * math.sqrt(3) * 128)
        f = vol.render(cam, 320, 240, brick_lod=lods)
        compare(f, oracle_frame(ds, pyr, lods, cam.as_dict(), 320, 240), f"persp cfg1 pose {k}")


@pytest.fixture(scope="module")
def cfg2(gpu_ctx):
    print("[gpu] generating config 2 for the marcher", flush=True)
    g = fcs.Generator(2)
    data = fcs.encode_volume(g.volume(), 64)
    vol = gpu_ctx.load(data)
    vol.decode_all(pyramid=True)
    ds = O.Dataset(data)
    pyr, _ = ds.decode_all_pyramid()
    return vol, ds, pyr


def test_lod_rule_matches_oracle(cfg2):
    vol, ds, _ = cfg2
    cam = vol.orbit_camera(math.pi / 4, 0.15 * math.pi, 1.5 * math.sqrt(3) * 512)
    lods = vol.brick_lods(cam, 1920, 1080)
    fov = 2 * math.atan(cam.s.tan_half_fov)
    for i in range(0, 512, 7):
        bx = sum(((i >> (3 * k)) & 1) << k for k in range(3))
        by = sum(((i >> (3 * k + 1)) & 1) << k for k in range(3))
        bz = sum(((i >> (3 * k + 2)) & 1) << k for k in range(3))
        c = (bx * 64 + 32.0, by * 64 + 32.0, bz * 64 + 32.0)
        want = O.brick_lod(c, 64.0, tuple(float(p) for p in cam.s.position), fov, 1080.0, 6)
        assert lods[i] == want


def test_orbit_poses_config2(cfg2):
    vol, ds, pyr = cfg2
    r = 1.5 * math.sqrt(3) * 512
    for k in range(8):
        cam = vol.orbit_camera(math.pi / 4 + k * math.pi / 4, 0.15 * math.pi, r)
        lods = vol.brick_lods(cam, 480, 270)
        f = vol.render(cam, 480, 270)
        assert np.array_equal(f.brick_lod_used, lods)
        compare(f, oracle_frame(ds, pyr, lods, cam.as_dict(), 480, 270), f"orbit cfg2 pose {k}")


def test_full_hd_frame_config2(cfg2):
    vol, ds, pyr = cfg2
    cam = vol.orbit_camera(math.pi / 4, 0.15 * math.pi, 1.5 * math.sqrt(3) * 512)
    f = vol.render(cam, 1920, 1080)
    compare(f, oracle_frame(ds, pyr, f.brick_lod_used, cam.as_dict(), 1920, 1080), "1080p cfg2")


def test_device_frames_against_exact_traversal(cfg2):
    """The device's own 1080p frames (every pose of the benchmark orbit, LOD
    rule) against the float64 exact traversal (oracle/dda_check.c): no pixel
    unexplained, grazing ties below 0.1 % of the pixels."""
    vol, ds, pyr = cfg2
    r = 1.5 * math.sqrt(3) * 512
    empty = [int(np.all(ds.palette(i) == 0)) for i in range(512)]
    p0 = [ds.palette(i)[0] for i in range(512)]
    tot = {"pixels": 0, "grazing": 0, "hits": 0}
    for k in range(8):
        cam = vol.orbit_camera(math.pi / 4 + k * math.pi / 4, 0.15 * math.pi, r)
        f = vol.render(cam, 1920, 1080)
        lods = f.brick_lod_used
        labels = [pyr[i][O.pyramid_offset(int(lods[i])):O.pyramid_offset(int(lods[i]) + 1)] for i in range(512)]
        st, _, _, _ = O.check_frame(8, 64, empty, lods, labels, p0, cam.as_dict(), 1920, 1080,
                                    f.label.astype(np.uint64), f.voxel, f.t_hit)
        print(f"[gpu] exact traversal pose {k}: {st}", flush=True)
        assert st["unexplained"] == 0, (k, st)
        for key in tot:
            tot[key] += st[key]
    assert tot["grazing"] <= 1e-3 * tot["pixels"], tot


def test_march_stats_match_column_walk(cfg1):
    """The march's read counts (samples, non-empty bricks entered, empty
    bricks tested) on an orthographic +z frame at the finest LOD equal a
    column walk: a ray tests the bricks of its column front to back until
    the brick of its first hit, samples every voxel of the non-empty ones up
    to the hit. The counts do not change the frame."""
    vol, ds, pyr = cfg1
    ctx = vol.ctx
    cam = fcs.Camera.ortho_plus_z(128)
    lods = np.full(64, 5, np.uint8)
    plain = vol.render(cam, 128, 128, brick_lod=lods)
    ctx.set_march_stats(True)
    try:
        f = vol.render(cam, 128, 128, brick_lod=lods)
        tm = ctx.timing()
    finally:
        ctx.set_march_stats(False)
    assert np.array_equal(f.label, plain.label) and np.array_equal(f.voxel, plain.voxel)
    vol_lab = np.zeros((128, 128, 128), np.uint64)  # z, y, x
    from conftest import c1
    idx = np.arange(32 ** 3)
    lx, ly, lz = c1(idx), c1(idx >> 1), c1(idx >> 2)
    empty = [bool(np.all(ds.palette(i) == 0)) for i in range(64)]
    for i in range(64):
        bx, by, bz = int(c1(i)), int(c1(i >> 1)), int(c1(i >> 2))
        vol_lab[bz * 32 + lz, by * 32 + ly, bx * 32 + lx] = pyr[i][O.pyramid_offset(5):]
    samples = bricks = tested_empty = 0
    for y in range(128):
        for x in range(128):
            for bz in range(4):
                bi = int(sum(((x // 32 >> k) & 1) << (3 * k) | ((y // 32 >> k) & 1) << (3 * k + 1) |
                             ((bz >> k) & 1) << (3 * k + 2) for k in range(3)))
                if empty[bi]:
                    tested_empty += 1
                    continue
                bricks += 1
                col = vol_lab[bz * 32:(bz + 1) * 32, y, x]
                nz = np.nonzero(col)[0]
                if len(nz):
                    samples += int(nz[0]) + 1
                    break
                samples += 32
    print(f"[gpu] march stats {tm['march_samples']} {tm['march_bricks_entered']} {tm['march_empty_tested']} "
          f"column walk {samples} {bricks} {tested_empty}", flush=True)
    assert (tm["march_samples"], tm["march_bricks_entered"], tm["march_empty_tested"]) == (samples, bricks, tested_empty)
