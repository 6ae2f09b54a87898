"""Learned brick prefetch around the on-demand renderer: visit-log collection
with the plain marcher, sample construction, training on the GPU
(include/fcs/fcs_mlp.h) and the evaluation of a path with and without the
network. The network only orders decodes; frames are compared by checksum."""
import time

import numpy as np

from .mlp import MLP, pack_bits
from .paths import extrapolate

N_POSE = 12


def collect_visits(vol, cams, w, h):
    """Render a path on demand from a cold cache; per frame the camera
    features and the visited-brick mask (the marcher's own log)."""
    vol.cache_clear()
    feats = np.zeros((len(cams), N_POSE), np.float32)
    vis = np.zeros((len(cams), vol.info["num_bricks"]), bool)
    for i, cam in enumerate(cams):
        fr = vol.render(cam, w, h, want=("brick_visited",))
        feats[i] = vol.camera_features(cam)
        vis[i] = fr.brick_visited != 0
    return feats, vis


def build_samples(feats, vis, horizon):
    """Sample f: (features_f, visited_f) -> the bricks the next `horizon`
    frames visit that frame f does not (the ones a prefetch can bring in;
    bricks visited now are already resident)."""
    n = len(feats) - 1
    tgt = np.zeros((n, vis.shape[1]), bool)
    for f in range(n):
        tgt[f] = vis[f + 1:f + 1 + horizon].any(0) & ~vis[f]
    return feats[:n], pack_bits(vis[:n]), pack_bits(tgt)


def train(ctx, nb, samples, hidden=512, epochs=8, batch=64, lr=1e-3, seed=0):
    poses = np.concatenate([s[0] for s in samples])
    inb = np.concatenate([s[1] for s in samples])
    tgt = np.concatenate([s[2] for s in samples])
    mlp = MLP(ctx, N_POSE + nb, hidden, nb, max_batch=batch, seed=seed)
    t0 = time.perf_counter()
    losses = mlp.train_masks(poses, inb, tgt, epochs=epochs, batch=batch, lr=lr, seed=seed + 1)
    return mlp, losses, time.perf_counter() - t0, len(poses)


def run_path(vol, cams, w, h, mlp=None, threshold=0.5, max_prefetch=2048, ahead=4, mode="plain", every=1):
    """Render a path from a cold cache: mode "plain" (fcs_render: frustum
    on-demand decode, the reference's form), "visible" (fcs_render_visible:
    decode the bricks rays enter) or with learned prefetch when `mlp` is given.
    Per frame: wall time, bricks decoded on demand, bricks prefetched, LOD
    used per brick, visited mask and the frame checksum."""
    vol.cache_clear()
    nf = len(cams)
    out = {"ms": np.zeros(nf), "device_ms": np.zeros(nf), "on_demand": np.zeros(nf, np.int64), "prefetched": [],
           "checksum": [], "visited": [], "lod": []}
    for i, cam in enumerate(cams):
        t0 = time.perf_counter()
        if mlp is None:
            if mode == "visible":
                fr = vol.render_visible(cam, w, h, want=("brick_visited",))
            else:
                fr = vol.render(cam, w, h, want=("brick_visited",))
            pf = np.zeros(0, np.uint32)
        elif i % every:
            # between network steps: the visibility-driven frame alone
            fr = vol.render_visible(cam, w, h, want=("brick_visited",))
            pf = np.zeros(0, np.uint32)
        else:
            cam_ahead = extrapolate(cams[i - 1], cam, ahead) if i > 0 else None
            fr, pf = vol.render_prefetch(mlp, cam, w, h, cam_ahead=cam_ahead, threshold=threshold,
                                         max_prefetch=max_prefetch, want=("brick_visited",))
        cs = vol.frame_checksum()
        out["ms"][i] = (time.perf_counter() - t0) * 1e3
        out["device_ms"][i] = fr.march_ms + fr.decode_ms
        out["on_demand"][i] = fr.bricks_decoded
        out["prefetched"].append(pf)
        out["checksum"].append(cs)
        out["visited"].append(fr.brick_visited != 0)
        out["lod"].append(fr.brick_lod_used.copy())
    return out


def wasted_prefetches(run):
    """Prefetched bricks never visited later in the path."""
    vis = np.array(run["visited"])
    nf = len(vis)
    later = np.zeros_like(vis)  # later[f] = visited in some frame after f
    for f in range(nf - 2, -1, -1):
        later[f] = later[f + 1] | vis[f + 1]
    total = wasted = 0
    for f, pf in enumerate(run["prefetched"]):
        if pf.size:
            total += pf.size
            wasted += int(np.count_nonzero(~later[f][pf]))
    return wasted, total


def compare(plain_runs, pf_runs):
    """Aggregate metrics over paths: frames/s of each form, miss rate (on-demand
    decodes left with prefetch / on-demand decodes of the plain form), wasted
    prefetches and frames whose checksum differs."""
    t_plain = sum(r["ms"].sum() for r in plain_runs)
    t_pf = sum(r["ms"].sum() for r in pf_runs)
    d_plain = sum(r["device_ms"].sum() for r in plain_runs)
    d_pf = sum(r["device_ms"].sum() for r in pf_runs)
    frames = sum(len(r["ms"]) for r in plain_runs)
    od_plain = sum(int(r["on_demand"].sum()) for r in plain_runs)
    od_pf = sum(int(r["on_demand"].sum()) for r in pf_runs)
    w = t = 0
    for r in pf_runs:
        a, b = wasted_prefetches(r)
        w += a
        t += b
    mism = sum(int(np.sum(np.array(a["checksum"], np.uint64) != np.array(b["checksum"], np.uint64)))
               for a, b in zip(plain_runs, pf_runs))
    return {
        "frames": frames,
        "fps_plain": frames / (t_plain * 1e-3),
        "fps_prefetch": frames / (t_pf * 1e-3),
        "speedup": t_plain / t_pf,
        "device_ms_per_frame_plain": d_plain / frames,
        "device_ms_per_frame_prefetch": d_pf / frames,
        "on_demand_plain": od_plain,
        "on_demand_prefetch": od_pf,
        "miss_rate": od_pf / max(od_plain, 1),
        "prefetched": t,
        "wasted": w,
        "wasted_frac": w / max(t, 1),
        "label_mismatch_frames": mism,
    }
