"""The learned-prefetch and serving workloads at their real sizes.

* Learned prefetch: the 1024^3 volume (seed 3), 32^3 bricks, a direct-mapped
  4 GiB cache, frames at 1920x1080 and the network at its real shape
  (12 + 32,768 -> 512 -> 512 -> 32,768, FP32), trained on the plain marcher's
  visit logs of training paths; on a held-out 600-frame fly-through the plain
  on-demand, visibility-driven and learned-prefetch frames have equal
  checksums, frame for frame (the network only orders decodes).
* Serving: the four serving volumes (256^3, 512^3, 1024^3, 2048^3, 32^3
  bricks) admitted whole under the 200 GB budget, 64 sessions, a Poisson
  trace at 2,000 frames/s over three resolutions on 8 views with replayed
  graphs: every served frame equals a single-frame render of its request.
"""
import math
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, file=sys.stderr, flush=True)


def test_learned_prefetch_1024_frames_identical(gpu_ctx):
    import fcs
    from fcs import paths
    from fcs import prefetch as PF

    say("[config4] encoding the 1024^3 volume")
    vol = gpu_ctx.load(fcs.encode_generated(fcs.Generator(4), 32))
    try:
        vol.cache_reserve_direct()
        nb, G = vol.info["num_bricks"], float(vol.info["grid_size"])
        assert nb == 32768 and vol.get_info()["cache_capacity_voxels"] == 32768 * 32768  # 4 GiB of u32 labels
        W, H = 1920, 1080
        samples = []
        for p in range(4):
            feats, vis = PF.collect_visits(vol, paths.flythrough(3, p, G, frames=200), W, H)
            samples.append(PF.build_samples(feats, vis, 12))
        mlp, losses, _, n = PF.train(gpu_ctx, nb, samples, hidden=512, epochs=4, batch=64, lr=1e-3)
        say(f"[config4] trained on {n} samples: loss {losses[0]:.4f} -> {losses[-1]:.4f}")
        assert losses[-1] < losses[0]
        cams = paths.flythrough(3, 20, G, frames=600)
        plain = PF.run_path(vol, cams, W, H, mode="plain")
        visible = PF.run_path(vol, cams, W, H, mode="visible")
        learned = PF.run_path(vol, cams, W, H, mlp=mlp, threshold=0.02, max_prefetch=4096, every=4, ahead=4)
        cv, cl = PF.compare([plain], [visible]), PF.compare([plain], [learned])
        say(f"[config4] visibility-driven {cv}\n[config4] learned {cl}")
        assert cv["label_mismatch_frames"] == 0 and cl["label_mismatch_frames"] == 0
        assert cl["prefetched"] > 0
        assert cv["on_demand_prefetch"] < cv["on_demand_plain"]
        mlp.close()
    finally:
        vol.close()


def test_serving_four_volumes_frames_equal_single_renders(gpu_ctx):
    import fcs
    from fcs import serve

    vols = []
    try:
        for s in (256, 512, 1024, 2048):
            say(f"[config5] encoding the {s}^3 serving volume")
            vols.append(gpu_ctx.load(fcs.encode_generated(fcs.Generator(5, size=s), 32)))
        plan, pool = serve.admit(vols, [1.0] * 4, 200 * 10 ** 9)
        assert list(plan) == [2, 2, 2, 2]  # every volume whole
        say(f"[config5] admitted {float(sum(int(p) for p in pool)) / 2 ** 30:.1f} GiB")
        views = [serve.View(gpu_ctx, 1920, 1080, graph=True, labels=False, checksum=True) for _ in range(8)]
        reqs, meta = serve.make_trace(vols, n_sessions=64, rate=2000.0, duration=0.25, seed=23)
        n = len(meta["arrival_s"])
        res = serve.run(views, vols, reqs)
        st = serve.latency_stats(meta, res)
        say(f"[config5] {n} requests served: {st}")
        assert st["missed_frames"] == 0
        bad = []
        for i in range(n):
            v = vols[int(meta["volume"][i])]
            cam = serve.request_camera(reqs[i])
            v.render(cam, int(meta["w"][i]), int(meta["h"][i]), want=())
            if v.frame_checksum() != int(res["checksum"][i]):
                bad.append(i)
        assert not bad, bad[:10]
        for vw in views:
            vw.close()
    finally:
        for v in vols:
            v.close()
