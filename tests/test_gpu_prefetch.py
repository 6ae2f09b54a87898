"""Direct-mapped cache, asynchronous prefetch, learned-prefetch frames and the
frame checksum, through the C ABI: decoded bricks bit-exact against the
oracle, frames identical with and without prefetch."""
import numpy as np
import pytest

import fcs
from conftest import golden_bytes
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def _cfg1(ctx):
    data = golden_bytes("cfg1_b32")
    return ctx.load(data, 32), O.Dataset(data, 32)


def test_direct_mapped_requests(gpu_ctx):
    print("[gpu] direct-mapped requests cfg1", flush=True)
    vol, ds = _cfg1(gpu_ctx)
    vol.cache_reserve_direct()
    rng = np.random.default_rng(7)
    bricks = rng.integers(0, 64, 300).astype(np.uint32)
    lods = rng.integers(0, 6, 300).astype(np.uint8)
    off, st = vol.request_bricks(bricks, lods)
    assert np.all(st == 0)
    assert np.all(off == bricks.astype(np.uint64) * 32768)
    last = {}
    for k in range(300):
        last[int(bricks[k])] = int(lods[k])
    for b, t in last.items():
        got = vol.read_cache(b * 32768, 1 << (3 * t)).astype(np.uint64)
        assert np.array_equal(got, ds.decode(b, t)[0]), (b, t)
    vol.close()


def test_prefetch_then_read(gpu_ctx):
    print("[gpu] prefetch cfg1", flush=True)
    vol, ds = _cfg1(gpu_ctx)
    vol.cache_reserve_direct()
    bricks = np.arange(64, dtype=np.uint32)
    lods = (np.arange(64) % 5 + 1).astype(np.uint8)
    n = vol.prefetch(bricks, lods)
    assert n == 64
    assert vol.prefetch(bricks, lods) == 0  # already resident
    for b in range(64):
        t = int(lods[b])
        got = vol.read_cache(b * 32768, 1 << (3 * t)).astype(np.uint64)
        assert np.array_equal(got, ds.decode(b, t)[0]), (b, t)
    vol.close()


def test_checksum_and_prefetch_frames_identical(gpu_ctx):
    print("[gpu] render with / without learned prefetch", flush=True)
    vol, _ = _cfg1(gpu_ctx)
    vol.cache_reserve_direct()
    nb = vol.info["num_bricks"]
    W, H = 160, 120
    cams = [vol.orbit_camera(0.3 + 0.05 * i, 0.2, 250.0) for i in range(6)]
    sums = []
    for cam in cams:
        vol.render(cam, W, H, want=())
        sums.append(vol.frame_checksum())
    vol.render(cams[0], W, H, want=())
    assert vol.frame_checksum() == sums[0]
    assert len(set(sums)) == len(sums)
    # learned prefetch (random network, threshold 0: prefetch everything it may)
    vol.cache_clear()
    mlp = fcs.MLP(gpu_ctx, 12 + nb, 32, nb, max_batch=4, seed=1)
    issued = 0
    for i, cam in enumerate(cams):
        ahead = cams[min(i + 1, len(cams) - 1)]
        _, pf = vol.render_prefetch(mlp, cam, W, H, cam_ahead=ahead, threshold=0.0, max_prefetch=nb, want=())
        issued += pf.size
        assert vol.frame_checksum() == sums[i], i
    assert issued > 0
    mlp.close()
    vol.close()


def test_camera_features(gpu_ctx):
    vol, _ = _cfg1(gpu_ctx)
    cam = vol.orbit_camera(0.5, 0.1, 300.0)
    f = vol.camera_features(cam)
    s = cam.as_dict()
    assert np.allclose(f[:3], np.array(s["position"]) / 128.0)
    assert np.allclose(f[3:6], s["forward"]) and np.allclose(f[6:9], s["up"]) and np.allclose(f[9:12], s["right"])
    vol.close()
