"""The Node.js module (N-API addon over the C ABI): loads under node and
exports the app-facing API (CPU); on a GPU, the JavaScript path decodes and
renders bit-exactly like the oracle."""
import json
import os
import shutil
import subprocess

import numpy as np
import pytest

from conftest import GOLDEN, ROOT, golden_bytes
from oracle import oracle as O

NODE = shutil.which("node")
ADDON = os.path.join(ROOT, "fast-compressed-segmentation_amd", "node", "fcs.node")
pytestmark = pytest.mark.skipif(NODE is None or not os.path.exists(ADDON), reason="node or the addon missing")

API = ["deviceCount", "createContext", "loadDatasetFromArrayBuffer", "decodeAllRansBricks", "decodeRansStatus",
       "readNibbles", "forceFullBrickRefresh", "reDecodeCachedBricks", "cacheResident", "cacheClear", "requestBricks",
       "decodeAll", "decodeVolume", "readBrick", "orbitCamera", "render", "updateLightingOptions", "applyLabelVisibility",
       "applyLabelHover", "getDatasetLabels", "transferFunction"]


def test_module_exports_app_api():
    out = subprocess.run([NODE, "-e", "console.log(JSON.stringify(Object.keys(require('"
                          + os.path.join(ROOT, "fast-compressed-segmentation_amd", "node") + "'))))"],
                         check=True, capture_output=True, text=True).stdout
    assert json.loads(out) == API


@pytest.mark.gpu
def test_node_path_bit_exact(tmp_path):
    import fcs

    if fcs.device_count() == 0:
        pytest.skip("no HIP device")
    print("[gpu] node decode_check.js on cfg1_b32", flush=True)
    res = subprocess.run([NODE, os.path.join(ROOT, "tests", "node", "decode_check.js"),
                          os.path.join(GOLDEN, "cfg1_b32.csdb"), str(tmp_path), "32"],
                         check=True, capture_output=True, text=True, timeout=120)
    meta = json.loads(res.stdout.strip().splitlines()[-1])
    assert meta["badMagicThrows"] is True
    st = np.fromfile(tmp_path / "rans_status.bin", np.uint32)
    assert st.shape == (64,) and np.all(st == 0)
    # decodeAllRansBricks resolved to one nibble Uint8Array per brick, equal to
    # the oracle's decodeRansStream restatement (decode.js:17-26)
    ds = O.Dataset(golden_bytes("cfg1_b32"), 32)
    nib = np.fromfile(tmp_path / "nibbles.bin", np.uint8)
    lens = np.fromfile(tmp_path / "nibble_lengths.bin", np.uint32)
    assert lens.shape == (64,) and meta["nibbleArrays"] == 64 and meta["nibblesAreUint8Arrays"] is True
    pos = 0
    for b in range(64):
        ref, rst = ds.rans(b)
        assert rst == 0 and int(lens[b]) == ref.size
        assert np.array_equal(nib[pos:pos + ref.size], ref), b
        pos += ref.size
    assert meta["readNibblesMatches"] is True
    # forceFullBrickRefresh / reDecodeCachedBricks decode the resident set again
    assert meta["refreshResident"] == meta["residentBefore"] > 0 and meta["refreshSameLabels"] is True
    vst = np.fromfile(tmp_path / "volume_status.bin", np.uint32)
    assert vst.shape == (64,) and np.all(vst == 0)
    assert np.array_equal(np.fromfile(tmp_path / "bricks_after_volume.bin", np.uint32),
                          np.fromfile(tmp_path / "bricks.bin", np.uint32))
    got = np.fromfile(tmp_path / "bricks.bin", np.uint32)
    pos = 0
    pyr, _ = ds.decode_all_pyramid()
    for b in range(64):
        for t in range(6):
            n = 1 << (3 * t)
            assert np.array_equal(got[pos:pos + n].astype(np.uint64), ds.decode(b, t)[0]), (b, t)
            pos += n
    lab = np.fromfile(tmp_path / "label.bin", np.uint32)
    vox = np.fromfile(tmp_path / "voxel.bin", np.float64)
    labels = [pyr[i][O.pyramid_offset(5):] for i in range(64)]
    empty = [int(np.all(ds.palette(i) == 0)) for i in range(64)]
    p0 = [ds.palette(i)[0] for i in range(64)]
    cam = {"position": [64.0, 64.0, -1.0], "forward": [0, 0, 1.0], "right": [1.0, 0, 0], "up": [0, -1.0, 0],
           "tan_half_fov": 0.0, "ortho_half_height": 64.0}
    rl, rv, _, _ = O.render(4, 32, empty, [5] * 64, labels, p0, cam, 256, 256)
    assert np.array_equal(lab.astype(np.uint64), rl)
    rvf = np.where(rv == np.uint64(2 ** 64 - 1), -1.0, rv.astype(np.float64))
    assert np.array_equal(vox, rvf)
    # shaded frame against the compositing oracle (same tolerance as the Python path)
    from test_gpu_shading import _compare_rgba, _tf_independent
    tl, tr = _tf_independent(ds)
    assert meta["numLabels"] == len(tl) and meta["hiddenAllTransparent"] is True
    a = 25.0 * np.pi / 180.0
    light = np.array([np.sin(a), 0.6, -np.cos(a)], np.float32)
    want = O.render_rgba(4, 32, empty, [5] * 64, labels, p0, cam, 256, 256, tl, tr, light)
    _compare_rgba(np.fromfile(tmp_path / "rgba.bin", np.uint32), want, "node shaded ortho")
