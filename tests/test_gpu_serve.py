"""Serving path on the device through include/fcs/fcs_serve.h: device-side
LOD selection against the host rule, view frames (eager and hipGraph
replays) against fcs_render, and the serving loop's frames against frames
rendered one at a time."""
import math

import numpy as np
import pytest

import fcs
from conftest import golden_bytes
from fcs import paths, serve
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def _volume(gpu_ctx, size, pyramid=True):
    data = fcs.encode_generated(fcs.Generator(5, size=size), 32)
    vol = gpu_ctx.load(data)
    if pyramid:
        vol.decode_all(pyramid=True)
    return vol, data


@pytest.fixture(scope="module")
def vol256(gpu_ctx):
    print("[gpu] serving volume 256^3, 32^3 bricks", flush=True)
    return _volume(gpu_ctx, 256)


def _cameras(vol):
    g = vol.info["grid_size"]
    diag = math.sqrt(3) * g
    cams = [vol.orbit_camera(0.3 + 0.7 * k, -0.4 + 0.2 * k, (0.8 + 0.15 * k) * diag) for k in range(5)]
    # inside the volume, looking along the diagonal and an axis
    cams.append(paths.camera_from((0.3 * g, 0.4 * g, 0.5 * g), (1.0, 0.7, 0.2)))
    cams.append(paths.camera_from((0.5 * g, 0.5 * g, 0.1 * g), (0.0, 0.0, 1.0)))
    cams.append(fcs.Camera.ortho_plus_z(g))
    return cams


def test_device_lod_rule_matches_host(gpu_ctx, vol256):
    vol, _ = vol256
    view = serve.View(gpu_ctx, 1920, 1080, graph=False, checksum=False)
    rng = np.random.default_rng(11)
    g = vol.info["grid_size"]
    cams = _cameras(vol)
    for k in range(24):  # random positions in and around the volume
        p = rng.uniform(-0.5 * g, 1.5 * g, 3)
        cams.append(paths.camera_from(p, rng.normal(size=3), fov=rng.uniform(0.3, 2.0)))
    for i, cam in enumerate(cams):
        for w, h in serve.RESOLUTIONS:
            print(f"[gpu] device LOD rule camera {i} {w}x{h}", flush=True)
            view.render_async(vol, cam, w, h)
            fr = view.read()
            assert np.array_equal(fr.brick_lod_used, vol.brick_lods(cam, w, h)), (i, w, h)
    view.close()


def _check_equal(fr, ref, tag):
    assert np.array_equal(fr.label, ref.label), tag
    assert np.array_equal(fr.voxel, ref.voxel), tag
    assert np.array_equal(fr.t_hit.view(np.uint32), ref.t_hit.view(np.uint32)), tag


@pytest.mark.parametrize("graph", [False, True])
def test_view_frames_equal_render(gpu_ctx, vol256, graph):
    vol, _ = vol256
    view = serve.View(gpu_ctx, 1920, 1080, graph=graph, labels=True, checksum=True)
    for i, cam in enumerate(_cameras(vol)):
        for w, h in ((640, 480), (1920, 1080)):
            print(f"[gpu] view frame graph={graph} camera {i} {w}x{h}", flush=True)
            ref = vol.render(cam, w, h, want=("label", "voxel", "t_hit"))
            assert ref.bricks_decoded == 0
            cs = vol.frame_checksum()
            view.render_async(vol, cam, w, h)
            r = view.wait()
            assert r["misses"] == 0 and r["checksum"] == cs, (i, w, h)
            assert np.array_equal(view.labels(), ref.label)
            _check_equal(view.read(), ref, (i, w, h))
    if graph:
        assert view.graphs == 2
    view.close()


def test_graph_replays_match_eager_frames(gpu_ctx, vol256):
    """Each (volume, size) graph is replayed many times with other cameras and
    with eager work of the program between the replays; every replay equals
    the eager view frame of the same camera."""
    vol, _ = vol256
    gview = serve.View(gpu_ctx, 1920, 1080, graph=True, labels=True, checksum=True)
    eview = serve.View(gpu_ctx, 1920, 1080, graph=False, labels=True, checksum=True)
    cams = _cameras(vol)
    order = [0, 1, 0, 2, 3, 0, 4, 5, 6, 7, 1, 1, 2]
    for n, i in enumerate(order):
        w, h = serve.RESOLUTIONS[n % 3]
        print(f"[gpu] graph replay {n}: camera {i} {w}x{h}", flush=True)
        eview.render_async(vol, cams[i], w, h)
        er = eview.wait()
        want = eview.labels()
        vol.render(cams[(i + 3) % len(cams)], 320, 240, want=("label",))  # eager work between replays
        gview.render_async(vol, cams[i], w, h)
        gr = gview.wait()
        assert gr["checksum"] == er["checksum"] and gr["misses"] == 0, (n, i)
        assert np.array_equal(gview.labels(), want), (n, i)
    assert gview.graphs == 3
    gview.close()
    eview.close()


def test_view_misses_and_fixed_lods(gpu_ctx, vol256):
    _, data = vol256
    vol = gpu_ctx.load(data)
    vol.decode_rans()
    nb = vol.info["num_bricks"]
    vol.cache_reserve(nb * 512)
    vol.request_bricks(np.arange(nb, dtype=np.uint32), np.full(nb, 3, np.uint8))
    view = serve.View(gpu_ctx, 640, 480, graph=False, checksum=True)
    cam = _cameras(vol)[1]
    fixed = np.full(nb, 3, np.uint8)
    view.render_async(vol, cam, 640, 480, brick_lod=fixed)
    r = view.wait()
    ref = vol.render(cam, 640, 480, brick_lod=fixed, want=("label", "voxel", "t_hit"))
    assert ref.bricks_decoded == 0
    assert r["misses"] == 0 and r["checksum"] == vol.frame_checksum()
    _check_equal(view.read(), ref, "fixed LOD 3")
    # the LOD rule asks for LOD 5 near the camera: those bricks are not resident
    view.render_async(vol, cam, 640, 480)
    assert view.wait()["misses"] > 0
    # rays pass through non-resident bricks: the frame equals the oracle's
    # with those bricks absent
    vol.cache_clear()
    keep = np.arange(0, nb, 3, dtype=np.uint32)
    vol.request_bricks(keep, np.full(keep.size, 3, np.uint8))
    view.render_async(vol, cam, 640, 480, brick_lod=fixed)
    assert view.wait()["misses"] > 0
    fr = view.read()
    ds = O.Dataset(data)
    pyr, _ = ds.decode_all_pyramid()
    B = round(nb ** (1 / 3))
    labels, empty, p0 = [], [], []
    for i in range(nb):
        pal = ds.palette(i)
        empty.append(int(np.all(pal == 0)))
        p0.append(pal[0])
        labels.append(pyr[i][O.pyramid_offset(3):O.pyramid_offset(4)] if i % 3 == 0 else None)
    lab, vox, tt, _ = O.render(B, 32, empty, fixed, labels, p0, cam.as_dict(), 640, 480)
    assert np.array_equal(fr.label.astype(np.uint64), lab) and np.array_equal(fr.voxel, vox)
    assert np.array_equal(fr.t_hit.view(np.uint32), tt.view(np.uint32))
    view.close()
    vol.close()


def _reference_checksums(vols_ref, reqs, n):
    out = np.zeros(n, np.uint64)
    for i in range(n):
        v = vols_ref[reqs[i].volume]
        v.render(serve.request_camera(reqs[i]), reqs[i].w, reqs[i].h, want=())
        out[i] = v.frame_checksum()
    return out


def test_serve_loop_frames_equal_single_frames(gpu_ctx, vol256):
    vol_b, _ = vol256
    vol_a, _ = _volume(gpu_ctx, 128, pyramid=False)
    vols = [vol_a, vol_b]
    plan, pool = serve.admit(vols, [1.0, 1.0], 10 * 2**30)
    assert list(plan) == [2, 2]
    views = [serve.View(gpu_ctx, 1920, 1080, graph=True, checksum=True) for _ in range(4)]
    reqs, meta = serve.make_trace(vols, n_sessions=8, rate=1000.0, duration=0.3, seed=5)
    n = len(meta["arrival_s"])
    print(f"[gpu] serving {n} requests on 4 views", flush=True)
    res = serve.run(views, vols, reqs)
    assert np.all(res["misses"] == 0)
    assert np.all(res["issue_s"] >= meta["arrival_s"] - 1e-9)
    assert np.all(res["done_s"] >= res["issue_s"])
    assert np.all(res["view"] < 4)
    want = _reference_checksums(vols, reqs, n)
    assert np.array_equal(res["checksum"], want)
    st = serve.latency_stats(meta, res)
    print(f"[gpu] p50 {st['p50_ms']:.3f} ms p99 {st['p99_ms']:.3f} ms", flush=True)
    # the same requests released at once: every view takes frames (how many
    # an open-loop trace keeps busy depends on arrival gaps vs frame time)
    for r in reqs:
        r.arrival_s = 0.0
    burst = serve.run(views, vols, reqs)
    assert len(set(burst["view"].tolist())) == 4
    assert np.array_equal(burst["checksum"], want)
    for v in views:
        v.close()
    vol_a.close()


def test_serve_loop_on_demand_fallback(gpu_ctx, vol256):
    """A volume admitted with an on-demand pool misses in the view frames; the
    loop serves those requests with fcs_render_visible and the frames equal
    those of a fully resident copy."""
    _, data256 = vol256
    vol_a, data128 = _volume(gpu_ctx, 128, pyramid=False)
    vol_b = gpu_ctx.load(data256)
    vols = [vol_a, vol_b]
    full = [serve.full_pyramid_bytes(v) for v in vols]
    plan, pool = serve.admit(vols, [1.0, 1.0], full[0] + full[1] - 4)
    assert list(plan) == [2, 1] and pool[1] == full[1] - 4
    views = [serve.View(gpu_ctx, 1920, 1080, graph=True, checksum=True) for _ in range(3)]
    reqs, meta = serve.make_trace(vols, n_sessions=4, rate=500.0, duration=0.2, seed=9,
                                  resolutions=((1920, 1080),), radius=(0.8, 1.2))
    n = len(meta["arrival_s"])
    print(f"[gpu] serving {n} requests with an on-demand pool", flush=True)
    res = serve.run(views, vols, reqs)
    assert np.any(res["misses"][meta["volume"] == 1] > 0)
    assert np.all(res["misses"][meta["volume"] == 0] == 0)
    ref_b, _ = _volume(gpu_ctx, 256)
    want = _reference_checksums([vol_a, ref_b], reqs, n)
    assert np.array_equal(res["checksum"], want)
    for v in views:
        v.close()
    for v in (vol_a, vol_b, ref_b):
        v.close()


def test_view_u64_labels(gpu_ctx, manifest):
    vol = gpu_ctx.load(golden_bytes("microns8_b64"), manifest["microns8_b64"]["brick_size"])
    assert vol.info["label_bits"] == 64
    vol.decode_all(pyramid=True)
    view = serve.View(gpu_ctx, 640, 480, graph=True, labels=True, checksum=True)
    for i, cam in enumerate(_cameras(vol)):
        ref = vol.render(cam, 640, 480, want=("label", "voxel", "t_hit"))
        cs = vol.frame_checksum()
        view.render_async(vol, cam, 640, 480)
        r = view.wait()
        assert r["label_bits"] == 64 and r["checksum"] == cs and r["misses"] == 0, i
        assert np.array_equal(view.labels(), ref.label)
        _check_equal(view.read(), ref, i)
    view.close()
    vol.close()


@pytest.mark.parametrize("graph", [False, True])
def test_shaded_view_frames_equal_render(gpu_ctx, vol256, graph):
    """Shaded views (the RGBA form) equal fcs_render's RGBA output; a lighting
    change is picked up by graph replays (a new graph is captured and the
    graph of the old lighting, which can never be replayed, is destroyed)."""
    vol, _ = vol256
    view = serve.View(gpu_ctx, 1280, 720, graph=graph, labels=True, checksum=True, rgba=True)
    cams = _cameras(vol)[:4]
    try:
        for angle in (25.0, -40.0):
            vol.set_lighting(angle, 0.5, 0.8)
            for i, cam in enumerate(cams):
                print(f"[gpu] shaded view graph={graph} light {angle} camera {i}", flush=True)
                ref = vol.render(cam, 640, 480, want=("label", "rgba"))
                view.render_async(vol, cam, 640, 480)
                r = view.wait()
                assert r["misses"] == 0
                assert np.array_equal(view.labels(), ref.rgba), (angle, i)
                assert np.array_equal(view.read_rgba(), ref.rgba)
                assert np.array_equal(view.read().label, ref.label)
            if graph:
                assert view.graphs == 1
    finally:
        vol.set_lighting()
        view.close()
