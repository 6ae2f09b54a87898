import json
import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "fast-compressed-segmentation_amd")
GOLDEN = os.path.join(ROOT, "tests", "golden")
for p in (ROOT, PKG):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a HIP device (MI355X)")


@pytest.fixture(scope="session")
def manifest():
    with open(os.path.join(GOLDEN, "manifest.json")) as f:
        return json.load(f)


def golden_bytes(name):
    with open(os.path.join(GOLDEN, name + ".csdb"), "rb") as f:
        return f.read()


def golden_volume(name, manifest):
    """Source volume of a golden case (config 1 is regenerated)."""
    if name == "cfg1_b32":
        import fcs

        return fcs.Generator(1).volume().astype(np.uint64)
    p = os.path.join(GOLDEN, name + "_volume.npy")
    if os.path.exists(p):
        return np.load(p, allow_pickle=False)
    return None


def c1(n):
    n = np.asarray(n, dtype=np.int64)
    n = n & 0x9249249
    n = (n ^ (n >> 2)) & 0x30C30C3
    n = (n ^ (n >> 4)) & 0x300F00F
    n = (n ^ (n >> 8)) & 0x30000FF
    n = (n ^ (n >> 16)) & 0x3FF
    return n


def brick_from_volume(vol, brick_id, b):
    """Finest-LOD labels of brick `brick_id` in Morton order (pad 0 outside)."""
    bx, by, bz = int(c1(brick_id)), int(c1(brick_id >> 1)), int(c1(brick_id >> 2))
    idx = np.arange(b ** 3)
    lx, ly, lz = c1(idx), c1(idx >> 1), c1(idx >> 2)
    D, H, W = vol.shape
    gx, gy, gz = bx * b + lx, by * b + ly, bz * b + lz
    inside = (gx < W) & (gy < H) & (gz < D)
    out = np.zeros(b ** 3, np.uint64)
    out[inside] = vol[gz[inside], gy[inside], gx[inside]]
    return out


def sha1(a):
    import hashlib

    return hashlib.sha1(np.ascontiguousarray(a).tobytes()).hexdigest()


@pytest.fixture(scope="session")
def gpu_ctx():
    import fcs

    if fcs.device_count() == 0:
        pytest.skip("no HIP device")
    ctx = fcs.Context(0)
    yield ctx
    ctx.close()
