"""Frames built on the device -- direct-mapped pools: the LOD rule, the
frustum test and the decode requests come from kernels, the op decoder drains
the device-built queue, and each frame is a replayed hipGraph -- against the
host-built frames of the same volume on a size-class pool (the LOD rule,
frustum test and request list on the host, renderLoop.js:268-298 and
cacheProcessing.js:99-232): every pixel's label, voxel index and t, the
visited mask, the LODs used, the number of bricks decoded and the resident
set equal, frame after frame along a moving camera, for the on-demand and the
visibility-driven forms, fixed LODs, shaded frames and u64 labels."""
import math

import numpy as np
import pytest

import fcs
from conftest import golden_bytes

pytestmark = pytest.mark.gpu

WANT = ("label", "voxel", "t_hit", "brick_visited")


@pytest.fixture(scope="module")
def pair(gpu_ctx):
    print("[gpu] device frames: 512^3 volume, 32^3 bricks", flush=True)
    data = fcs.encode_generated(fcs.Generator(5, size=512), 32)
    vd, vh = gpu_ctx.load(data), gpu_ctx.load(data)
    vd.cache_reserve_direct()
    yield vd, vh
    vd.close()
    vh.close()


def cams(vol, n, radius=1.2):
    G = vol.info["grid_size"]
    return [vol.orbit_camera(0.4 + 0.09 * k, 0.25 - 0.03 * k, (radius - 0.05 * k) * math.sqrt(3) * G)
            for k in range(n)]


def same(fd, fh, tag):
    assert np.array_equal(fd.label, fh.label), tag
    assert np.array_equal(fd.voxel, fh.voxel), tag
    assert np.array_equal(fd.t_hit.view(np.uint32), fh.t_hit.view(np.uint32)), tag
    assert np.array_equal(fd.brick_visited, fh.brick_visited), tag
    assert np.array_equal(fd.brick_lod_used, fh.brick_lod_used), tag
    assert fd.bricks_decoded == fh.bricks_decoded, (tag, fd.bricks_decoded, fh.bricks_decoded)


@pytest.mark.parametrize("form", ["render", "render_visible"])
def test_device_frames_equal_host_frames(pair, form):
    vd, vh = pair
    for v in (vd, vh):
        v.cache_clear()
    decoded = 0
    for k, cam in enumerate(cams(vd, 10)):
        print(f"[gpu] {form} frame {k}", flush=True)
        fd = getattr(vd, form)(cam, 640, 360, want=WANT)
        fh = getattr(vh, form)(cam, 640, 360, want=WANT)
        same(fd, fh, (form, k))
        decoded += fd.bricks_decoded
        if k % 3 == 2:
            bd, td = vd.cache_resident()
            bh, th = vh.cache_resident()
            assert np.array_equal(bd, bh) and np.array_equal(td, th), (form, k)
    assert decoded > 0
    # the host sees what device frames decoded: requesting it again decodes nothing new
    bd, td = vd.cache_resident()
    off, st = vd.request_bricks(bd, td)
    assert not np.any(st) and np.array_equal(off, bd.astype(np.uint64) * np.uint64(32768))


def test_device_frames_fixed_lods_and_shading(pair):
    vd, vh = pair
    for v in (vd, vh):
        v.cache_clear()
    rng = np.random.default_rng(3)
    lods = rng.integers(0, 6, vd.info["num_bricks"]).astype(np.uint8)
    for k, cam in enumerate(cams(vd, 4, 1.0)):
        fd = vd.render(cam, 480, 270, brick_lod=lods, want=WANT + ("rgba",))
        fh = vh.render(cam, 480, 270, brick_lod=lods, want=WANT + ("rgba",))
        same(fd, fh, ("fixed+rgba", k))
        assert np.array_equal(fd.rgba, fh.rgba), k
    # a lighting change reaches the replayed frame
    for v in (vd, vh):
        v.set_lighting(-30.0, 0.4, 0.9)
    cam = cams(vd, 1, 1.0)[0]
    fd = vd.render(cam, 480, 270, brick_lod=lods, want=("rgba",))
    fh = vh.render(cam, 480, 270, brick_lod=lods, want=("rgba",))
    assert np.array_equal(fd.rgba, fh.rgba)
    for v in (vd, vh):
        v.set_lighting()


def test_device_frames_u64_labels(gpu_ctx):
    data = golden_bytes("u64_b8")
    vd, vh = gpu_ctx.load(data, 8), gpu_ctx.load(data, 8)
    try:
        assert vd.info["label_bits"] == 64
        vd.cache_reserve_direct()
        G = vd.info["grid_size"]
        for k in range(3):
            cam = vd.orbit_camera(0.5 + k, 0.2, 1.5 * math.sqrt(3) * G)
            same(vd.render_visible(cam, 200, 150, want=WANT), vh.render_visible(cam, 200, 150, want=WANT), k)
    finally:
        vd.close()
        vh.close()
