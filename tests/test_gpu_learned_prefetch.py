"""Learned prefetch end to end at small scale: visit logs from the plain
marcher train the network on the GPU; on held-out fly-through paths the
frames are identical (checksums) to plain on-demand rendering, and the trained
network removes on-demand decodes that the visibility-driven form without it
still pays."""
import numpy as np
import pytest

import fcs
from fcs import paths
from fcs import prefetch as PF

pytestmark = pytest.mark.gpu


def test_learned_prefetch_gains_over_on_demand(gpu_ctx):
    print("[gpu] learned prefetch, 256^3 volume, 32^3 bricks", flush=True)
    gen = fcs.Generator(4, size=256)
    vol = gpu_ctx.load(fcs.encode_generated(gen, 32))
    vol.cache_reserve_direct()
    nb, G = vol.info["num_bricks"], float(vol.info["grid_size"])
    W, H, F = 320, 240, 150
    samples = []
    for p in range(10):
        feats, vis = PF.collect_visits(vol, paths.flythrough(7, p, G, frames=F), W, H)
        samples.append(PF.build_samples(feats, vis, 8))
    mlp, losses, _, _ = PF.train(gpu_ctx, nb, samples, hidden=128, epochs=30, batch=32, lr=2e-3, seed=3)
    assert losses[-1] < losses[0]
    plain, visible, learned = [], [], []
    for p in range(10, 13):
        cams = paths.flythrough(7, p, G, frames=F)
        plain.append(PF.run_path(vol, cams, W, H, mode="plain"))
        visible.append(PF.run_path(vol, cams, W, H, mode="visible"))
        learned.append(PF.run_path(vol, cams, W, H, mlp=mlp, threshold=0.05, max_prefetch=nb, every=2, ahead=2))
    cv = PF.compare(plain, visible)
    cl = PF.compare(plain, learned)
    print("visible", cv, "\nlearned", cl, flush=True)
    assert cv["label_mismatch_frames"] == 0 and cl["label_mismatch_frames"] == 0
    assert cl["prefetched"] > 0
    assert cv["on_demand_prefetch"] < cv["on_demand_plain"]
    assert cl["on_demand_prefetch"] < cv["on_demand_prefetch"]
    mlp.close()
    vol.close()
