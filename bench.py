#!/usr/bin/env python3
"""Benchmark: brick decoder + DDA marcher on one MI355X.

Default workload (--config 3, the headline): the 2048^3 uint32 regions volume
(seed 2), 32^3 bricks (262,144), encoded to .csdb by the CPU producer and
resident in HBM before timing. One step = the rANS stage + the decode of every
brick at its finest LOD (8.59 G labels), both in one call
(fcs_decode_volume). value = decoded Gvoxels/s (HIP events on the library
stream). Every label of the last step is then compared with the source
volume, a batched launch of 100k (brick, LOD) requests is timed and compared
with the oracle, the march over the decoded volume is timed at 1080p, and the
1024^3 fly-throughs of --config 4 are run for the metric's chosen-vs-plain
frames/s (learned prefetch against the plain on-demand form), which the line
carries as `chosen_vs_plain_1080p`.

Other workloads: --config 2 (512^3 demo scale, 64^3 bricks: full pyramid
decode + 8 orbit poses at 1080p), --config 4 (1024^3 fly-throughs, learned
prefetch vs on-demand), --config 5 (serving many viewers). Prints one JSON
line.
"""
import argparse
import json
import math
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "fast-compressed-segmentation_amd"))

HBM_PEAK_GBS = 8000.0
RANS_MAPPINGS = {1: "lane per brick", 2: "wave per brick (scalar table)", 3: "wave per brick (ballot)",
                 4: "wave per brick (LDS table)", 5: "wave per brick, scalar byte window (table)",
                 6: "wave per brick, scalar byte window (ballot)"}


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def measured_traffic(summary, kernel):
    """HBM bytes per launch of `kernel` from a committed counter summary
    (profiles/*.md, lines 'traffic_bytes: <kernel> <bytes>'), or None."""
    path = os.path.join(ROOT, "profiles", summary)
    try:
        with open(path) as f:
            for line in f:
                if line.startswith("traffic_bytes:"):
                    _, k, b = line.split()
                    if k == kernel:
                        return int(b)
    except OSError:
        pass
    return None


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-config4", action="store_true",
                    help="config 3: leave out the learned-prefetch fly-throughs (config 4) from the line")
    ap.add_argument("--cpu-seconds", type=float, default=12.0)
    ap.add_argument("--config", type=int, default=3, choices=(2, 3, 4, 5),
                    help="3: 2048^3 full decode + 100k batched requests (default); 2: 512^3 demo scale; "
                         "4: 1024^3 fly-throughs, learned prefetch vs on-demand; "
                         "5: serving 64 viewers over 256^3..2048^3 volumes")
    ap.add_argument("--size", type=int, default=0, help="override the volume edge (config 3)")
    ap.add_argument("--train-paths", type=int, default=20)
    ap.add_argument("--eval-paths", type=int, default=5)
    ap.add_argument("--frames", type=int, default=600)
    ap.add_argument("--epochs", type=int, default=16)
    ap.add_argument("--threshold", type=float, default=0.02)
    ap.add_argument("--every", type=int, default=4)
    ap.add_argument("--sizes", type=str, default="256,512,1024,2048", help="serving volume edges (config 5)")
    ap.add_argument("--sessions", type=int, default=64)
    ap.add_argument("--rate", type=float, default=2000.0, help="offered frames/s (config 5)")
    ap.add_argument("--duration", type=float, default=5.0, help="seconds of traffic (config 5)")
    ap.add_argument("--views", type=int, default=8, help="concurrent render targets (config 5)")
    ap.add_argument("--budget-gb", type=float, default=200.0)
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="BYTES",
                    help="only run 3 device-to-device copies of BYTES (HBM counter calibration)")
    args = ap.parse_args()
    if args.copy_calibration:
        return copy_calibration(args.copy_calibration)
    if args.gpus != 1:
        print(json.dumps({"error": "this benchmark runs on exactly one GPU; sharding the volume across "
                                   "GPUs is not built", "n_gpus": args.gpus}))
        sys.exit(2)

    import numpy as np

    import fcs

    if args.config == 3:
        return bench_config3(args)
    if args.config == 4:
        return bench_config4(args)
    if args.config == 5:
        return bench_config5(args)

    t0 = time.time()
    gen = fcs.Generator(2)
    data = fcs.encode_generated(gen, 64)
    log(f"[bench] generated + encoded 512^3 volume: {len(data)} bytes in {time.time() - t0:.1f}s")
    ctx = fcs.Context(0)
    vol = ctx.load(data)
    info = vol.info
    G = info["grid_size"]
    radius = 1.5 * math.sqrt(3.0) * G
    cams = [vol.orbit_camera(math.pi / 4 + k * math.pi / 4, 0.15 * math.pi, radius, math.pi / 3) for k in range(8)]
    W, H = 1920, 1080

    def step():
        st = vol.decode_rans()
        if np.any(st):
            raise RuntimeError("corrupt bricks in the rANS stage")
        tr = ctx.timing()
        vol.decode_all(pyramid=True)
        td = ctx.timing()
        marches = []
        for cam in cams:
            f = vol.render(cam, W, H, want=())
            if f.bricks_decoded:
                raise RuntimeError("frame needed a decode after the full decode")
            marches.append(f.march_ms)
        return tr["rans_ms"], td["decode_ms"], marches, td["decoded_voxels"], td["decode_bytes"]

    first = vol.decode_rans()  # autotunes the rANS mapping on this volume
    tune = ctx.timing()
    for _ in range(args.warmup):
        step()
    ctx.sync()
    recs = []
    t1 = time.perf_counter()
    for _ in range(args.steps):
        recs.append(step())
    ctx.sync()
    wall = time.perf_counter() - t1

    rans_ms = float(np.mean([r[0] for r in recs]))
    ops_ms = float(np.mean([r[1] for r in recs]))
    march_ms = float(np.mean([np.mean(r[2]) for r in recs]))
    voxels = recs[0][3]
    ops_bytes = recs[0][4]
    decode_ms = rans_ms + ops_ms
    gvox = voxels / (decode_ms * 1e-3) / 1e9
    fps = 1000.0 / march_ms
    # rANS stage algorithmic bytes: compressed bytes read + nibble bytes
    # written + 8 B per level entry + 32 B brick record
    rans_bytes = info["compressed_bytes"] + info["total_symbols"] // 2 + info["num_bricks"] * (8 * (info["levels"] + 1) + 32)
    kernels = {
        "rans_kernel": {"ms": rans_ms, "bytes": rans_bytes, "gbs": rans_bytes / (rans_ms * 1e-3) / 1e9},
        "ops_kernel": {"ms": ops_ms, "bytes": ops_bytes, "gbs": ops_bytes / (ops_ms * 1e-3) / 1e9},
        "march_kernel": {"ms": march_ms, "bytes": None, "gbs": None},
    }
    ctx.set_march_stats(True)
    mb = []
    for cam in cams:
        vol.render(cam, W, H, want=())
        tm = ctx.timing()
        mb.append(4 * tm["march_samples"] + 48 * tm["march_bricks_entered"] + 32 * tm["march_empty_tested"]
                  + W * H * 16)
    ctx.set_march_stats(False)
    kernels["march_kernel"]["bytes"] = float(np.mean(mb))
    kernels["march_kernel"]["gbs"] = kernels["march_kernel"]["bytes"] / (march_ms * 1e-3) / 1e9
    dom = max(("rans_kernel", "ops_kernel"), key=lambda k: kernels[k]["ms"])
    if march_ms > kernels[dom]["ms"]:
        pass  # the march is reported per frame; decode kernels carry the roofline
    ach = kernels[dom]["gbs"]
    out = {
        "metric": "decoded Gvoxels/s + 1080p frames/s, chosen vs plain GPU form, bit-exact",
        "value": gvox,
        "unit": "Gvoxels/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": wall * 1000.0 / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u32",
        "data": "synthetic",
        "config": {"workload": "512^3 u32 Voronoi (20k cells, seed 1), 64^3 bricks x512: full decode of every "
                               "LOD + 8 orbit poses at 1920x1080", "bricks": info["num_bricks"],
                   "brick_size": info["brick_size"], "compressed_bytes": info["compressed_bytes"],
                   "symbols": info["total_symbols"], "form": "plain on-demand"},
        "frames_per_s_1080p": fps,
        "rays_per_s": W * H * fps,
        "decode_ms": decode_ms,
        "rans_ms": rans_ms,
        "ops_ms": ops_ms,
        "march_ms_per_frame": march_ms,
        "decoded_voxels_per_step": voxels,
        "roofline": {"bound": "hbm", "kernel": dom, "achieved": ach, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": ach / HBM_PEAK_GBS,
                     "traffic": measured_traffic("hbm_traffic_512cube_b64.md", dom)},
        "kernels": kernels,
        "rans_autotune": {"mapping": RANS_MAPPINGS.get(tune["rans_mapping"]),
                          "lane_ms": tune["tune_lane_ms"], "wave_table_ms": tune["tune_wave_ms"],
                          "wave_ballot_ms": tune["tune_ballot_ms"], "wave_lds_ms": tune["tune_wave_lds_ms"],
                          "wave_window_table_ms": tune["tune_win_ms"],
                          "wave_window_ballot_ms": tune["tune_win_ballot_ms"]},
    }
    if not args.no_cpu_baseline:
        from oracle import oracle as O

        threads = os.cpu_count() or 1
        ds = O.Dataset(data)
        t2 = time.time()
        reps, vox = 0, 0
        pyr = None
        while reps == 0 or time.time() - t2 < args.cpu_seconds * 0.6:
            pyr, _ = ds.decode_all_pyramid(threads)
            vox += pyr.size
            reps += 1
        dec_s = time.time() - t2
        cam = cams[0]
        lods = vol.brick_lods(cam, W, H)
        labels = [pyr[i][O.pyramid_offset(int(lods[i])):O.pyramid_offset(int(lods[i]) + 1)] for i in range(512)]
        empty = [int(np.all(ds.palette(i) == 0)) for i in range(512)]
        p0 = [ds.palette(i)[0] for i in range(512)]
        t3 = time.time()
        O.render(info["bricks_per_axis"], info["brick_size"], empty, lods, labels, p0, cam.as_dict(), W, H)
        frame_s = time.time() - t3
        out["cpu_baseline"] = {
            "value": vox / dec_s / 1e9, "unit": "Gvoxels/s", "cores": threads, "kind": "port",
            "sample": f"oracle full decode (every LOD of all 512 bricks) x{reps} on {threads} threads; "
                      f"one 1080p frame single-threaded: {1.0 / frame_s:.3f} frames/s",
            "frames_per_s_1080p_1thread": 1.0 / frame_s,
        }
    print(json.dumps(out))


def copy_calibration(nbytes):
    """Three device-to-device copies of `nbytes` (a plain copy kernel of known
    size): the FETCH_SIZE / WRITE_SIZE counters of one such dispatch against
    nbytes give the scale factor for the decoder's counter readings."""
    import ctypes

    hip = ctypes.CDLL("libamdhip64.so")
    a, b = ctypes.c_void_p(), ctypes.c_void_p()
    n = ctypes.c_size_t(nbytes)
    if hip.hipMalloc(ctypes.byref(a), n) or hip.hipMalloc(ctypes.byref(b), n):
        raise RuntimeError("hipMalloc failed")
    hip.hipMemset(a, 1, n)
    for _ in range(3):
        if hip.hipMemcpy(b, a, n, 3):  # hipMemcpyDeviceToDevice
            raise RuntimeError("hipMemcpy failed")
    hip.hipDeviceSynchronize()
    hip.hipFree(a)
    hip.hipFree(b)
    print(json.dumps({"copy_bytes": nbytes, "copies": 3}))


def bench_config3(args):
    """The headline workload: the 2048^3 u32 regions volume (seed 2, 256^3
    regions of background / ~50 cells / ~4000 noisy cells), 32^3 bricks
    (262,144), encoded by the CPU producer and resident in HBM before timing.

    One step = the rANS stage + the decode of every brick at its finest LOD
    (8.59 G labels, 34.4 GB), both launches back to back in one call
    (fcs_decode_volume). value = decoded Gvoxels/s of that call (HIP events on
    the library stream). After the timed steps every label of the last decode
    is compared with the generator's source bricks (the finest LOD equals the
    source volume), and one batched launch of 100k shuffled (brick, LOD uniform
    in 0..5) requests is timed and compared with the oracle. Each stage is
    also timed on its own for the roofline (the op stage is the dominant
    kernel). frames_per_s_1080p: the march over the decoded volume at 1920x1080
    for 8 orbit poses (every brick at its finest LOD, so no frame decodes)."""
    import numpy as np

    import fcs

    t0 = time.time()
    gen = fcs.Generator(3, size=args.size or 0)
    data = fcs.encode_generated(gen, 32)
    log(f"[bench] generated + encoded {gen.size}^3 regions volume: {len(data)} bytes in {time.time() - t0:.1f}s")
    ctx = fcs.Context(0)
    vol = ctx.load(data)
    info = vol.info
    nb, L = info["num_bricks"], info["levels"]
    per = 1 << (3 * L)
    st = vol.decode_rans()
    if np.any(st):
        raise RuntimeError("corrupt bricks")
    tune = ctx.timing()
    vol.cache_reserve(nb * per)
    vol.decode_all(lod=L, pyramid=False)  # the op decoder's split autotune runs on this first full batch

    def step():
        st = vol.decode_volume(lod=L, pyramid=False)
        if np.any(st):
            raise RuntimeError("corrupt bricks")
        td = ctx.timing()
        return td["pipeline_ms"], td["decoded_voxels"], td["decode_bytes"]

    def stages():
        vol.decode_rans()
        tr = ctx.timing()["rans_ms"]
        vol.decode_all(lod=L, pyramid=False)
        return tr, ctx.timing()["decode_ms"]

    for _ in range(args.warmup):
        step()
    ctx.sync()
    recs = []
    t1 = time.perf_counter()
    for _ in range(args.steps):
        recs.append(step())
    ctx.sync()
    wall = time.perf_counter() - t1
    log(f"[bench] {args.steps} timed steps: {wall * 1000.0 / args.steps:.2f} ms per step")
    dec_ms = float(np.mean([r[0] for r in recs]))
    voxels, ops_bytes = recs[0][1], recs[0][2]

    # every label of the last timed decode against the source volume
    t2 = time.time()
    off, st2 = vol.request_bricks(np.arange(nb, dtype=np.uint32), np.full(nb, L, np.uint8))
    if np.any(st2) or not np.array_equal(off, np.arange(nb, dtype=np.uint64) * np.uint64(per)):
        raise RuntimeError("decoded bricks are not resident in slot order")
    bad_bricks = 0
    chunk = 8192
    for first in range(0, nb, chunk):
        n = min(chunk, nb - first)
        dev = vol.read_cache(first * per, n * per).reshape(n, per)
        src = gen.bricks(32, first, n)
        bad_bricks += int(np.count_nonzero(np.any(dev != src, axis=1)))
    check_s = time.time() - t2
    log(f"[bench] labels checked against the source volume in {check_s:.1f}s: {bad_bricks} bricks differ")

    srec = [stages() for _ in range(max(1, args.steps))]
    rans_ms = float(np.mean([r[0] for r in srec]))
    ops_ms = float(np.mean([r[1] for r in srec]))
    rans_bytes = info["compressed_bytes"] + info["total_symbols"] // 2 + nb * 32

    # the march over the decoded volume: 8 orbit poses at 1080p, finest LOD
    W, H = 1920, 1080
    G = info["grid_size"]
    radius = 1.5 * math.sqrt(3.0) * G
    cams = [vol.orbit_camera(math.pi / 4 + k * math.pi / 4, 0.15 * math.pi, radius, math.pi / 3) for k in range(8)]
    finest = np.full(nb, L, np.uint8)
    vol.render(cams[0], W, H, brick_lod=finest, want=())
    march = []
    for cam in cams:
        f = vol.render(cam, W, H, brick_lod=finest, want=())
        if f.bricks_decoded:
            raise RuntimeError("frame needed a decode after the full decode")
        march.append(f.march_ms)
    march_ms = float(np.mean(march))
    # the march's algorithmic bytes per frame: what the rays read (4 B per
    # label sample, the 48 B records of each non-empty brick entered, 32 B per
    # empty brick tested) and the per-pixel outputs (label 4 B, voxel index
    # 8 B, t 4 B), counted by a separate pass over the same poses
    ctx.set_march_stats(True)
    mbytes = []
    for cam in cams:
        vol.render(cam, W, H, brick_lod=finest, want=())
        tm = ctx.timing()
        mbytes.append(4 * tm["march_samples"] + 48 * tm["march_bricks_entered"] + 32 * tm["march_empty_tested"]
                      + W * H * 16)
    ctx.set_march_stats(False)
    march_bytes = float(np.mean(mbytes))
    from oracle import oracle as O

    rng = np.random.default_rng(2)
    nreq = min(100000, nb)
    bricks = rng.choice(nb, nreq, replace=False).astype(np.uint32)
    lods = rng.integers(0, L + 1, nreq).astype(np.uint8)
    vol.cache_clear()
    boff, bst = vol.request_bricks(bricks, lods)
    tb = ctx.timing()
    bvox, bbytes = tb["decoded_voxels"], tb["decode_bytes"]
    sizes = np.left_shift(np.uint64(1), np.uint64(3) * lods.astype(np.uint64))
    bdev = vol.read_cache(0, int((boff + sizes).max()))
    ds = O.Dataset(data)
    threads = os.cpu_count() or 1
    bref, broff, brst = ds.decode_requests(bricks, lods, np.uint32, threads)
    bbad = sum(1 for k in range(nreq)
               if not np.array_equal(bdev[int(boff[k]):int(boff[k] + sizes[k])],
                                     bref[int(broff[k]):int(broff[k] + sizes[k])]))
    del bdev, bref
    bms = []
    for _ in range(3):
        vol.cache_clear()
        vol.request_bricks(bricks, lods)
        bms.append(ctx.timing()["decode_ms"])
    batch_ms = float(np.median(bms))
    tops = ctx.timing()
    ach = ops_bytes / (ops_ms * 1e-3) / 1e9
    pipe_ach = ops_bytes / (dec_ms * 1e-3) / 1e9
    exact = bad_bricks == 0 and bbad == 0 and not np.any(bst)
    out = {
        "metric": "decoded Gvoxels/s + 1080p frames/s, chosen vs plain GPU form, bit-exact",
        "value": voxels / (dec_ms * 1e-3) / 1e9,
        "unit": "Gvoxels/s",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": wall * 1000.0 / args.steps,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "u32", "data": "synthetic",
        "config": {"workload": f"{gen.size}^3 u32 regions volume (seed 2), 32^3 bricks x{nb}: rANS + finest-LOD "
                               f"decode of every brick (fcs_decode_volume: both stages in one call)",
                   "compressed_bytes": info["compressed_bytes"], "symbols": info["total_symbols"],
                   "form": "full decode"},
        "bit_exact": bool(exact),
        "label_check": {"labels": int(nb) * per, "bricks_differing_from_source": bad_bricks,
                        "batched_requests_differing_from_oracle": int(bbad), "seconds": check_s},
        "decode_ms": dec_ms, "decode_ms_steps": [round(r[0], 3) for r in recs], "rans_ms": rans_ms, "ops_ms": ops_ms,
        "decoded_voxels_per_step": int(voxels),
        "two_calls_gvox_s": voxels / ((rans_ms + ops_ms) * 1e-3) / 1e9,
        "ops_only_gvox_s": voxels / (ops_ms * 1e-3) / 1e9,
        "decode_call_gbs": pipe_ach, "decode_call_frac": pipe_ach / HBM_PEAK_GBS,
        # the dominant kernel: the op stage (head pass + op kernel), timed by
        # the library's events on the stream it runs on
        "roofline": {"bound": "hbm", "kernel": "ops_head_kernel+ops_kernel", "achieved": ach, "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": ach / HBM_PEAK_GBS,
                     "traffic": measured_traffic("hbm_traffic_2048cube_b32.md", "ops_stage")
                     if gen.size == 2048 else None,
                     "traffic_source": "rocprofv3 FETCH_SIZE + WRITE_SIZE passes, profiles/hbm_traffic_2048cube_b32.md"},
        "frames_per_s_1080p": 1000.0 / march_ms, "rays_per_s": W * H * 1000.0 / march_ms,
        "march_ms_per_frame": march_ms,
        "march_kernel": {"ms": march_ms, "bytes_per_frame": march_bytes,
                         "gbs": march_bytes / (march_ms * 1e-3) / 1e9,
                         "frac": march_bytes / (march_ms * 1e-3) / 1e9 / HBM_PEAK_GBS,
                         "traffic": measured_traffic("hbm_traffic_2048cube_b32.md", "march_kernel")
                         if gen.size == 2048 else None},
        "rans_gbs": rans_bytes / (rans_ms * 1e-3) / 1e9,
        "rans_autotune": {"mapping": RANS_MAPPINGS.get(tune["rans_mapping"]), "lane_ms": tune["tune_lane_ms"],
                          "wave_table_ms": tune["tune_wave_ms"], "wave_ballot_ms": tune["tune_ballot_ms"],
                          "wave_lds_ms": tune["tune_wave_lds_ms"], "wave_window_table_ms": tune["tune_win_ms"],
                          "wave_window_ballot_ms": tune["tune_win_ballot_ms"]},
        "ops_autotune": {"split": {1: "3 workgroups/CU", 2: "4 workgroups/CU"}.get(tops["ops_split"]),
                         "3wg_ms": tops["tune_ops_3wg_ms"], "4wg_ms": tops["tune_ops_4wg_ms"]},
        "batched": {"requests": int(nreq), "lods": "uniform 0..%d" % L, "ms": batch_ms,
                    "gvox_s": bvox / (batch_ms * 1e-3) / 1e9, "gbs": bbytes / (batch_ms * 1e-3) / 1e9,
                    "status_ok": bool(np.all(bst == 0)), "differing_from_oracle": int(bbad)},
    }
    if not args.no_cpu_baseline:
        out["cpu_baseline"] = cpu_baseline_config3(args, data, ds, nb, L)
    if not args.no_config4:
        # the metric's "chosen vs plain GPU form": learned prefetch against the
        # plain on-demand form on the 1024^3 fly-throughs (config 4, full size)
        vol.close()
        ctx.close()
        c4 = run_config4(args)
        out["chosen_vs_plain_1080p"] = {
            "workload": c4["config"]["workload"],
            "fps_plain_on_demand": c4["fps_plain_on_demand"],
            "fps_visibility_driven": c4["fps_visibility_driven"],
            "fps_learned_prefetch": c4["fps_learned_prefetch"],
            "speedup_learned_vs_plain": c4["speedup_vs_plain"],
            "speedup_visibility_driven_vs_plain": c4["speedup_visibility_driven_vs_plain"],
            "device_ms_per_frame": c4["device_ms_per_frame"],
            "on_demand_decodes": c4["on_demand_decodes"],
            "miss_rate": c4["miss_rate"], "wasted_frac": c4["wasted_frac"],
            "label_mismatch_frames": c4["label_mismatch_frames"],
        }
    print(json.dumps(out))


def cpu_baseline_config3(args, data, ds, nb, L):
    """Finest-LOD decode of random bricks of the volume on every host core,
    each for about cpu_seconds / 2: the oracle restatement ("port") and, where
    oracle/_ref was built, the reference's own decodeRansStream +
    decodeBrick<32> ("reference", over non-constant bricks: the reference
    decoder rejects a brick whose root stop bit is set)."""
    import ctypes
    import tempfile
    from concurrent.futures import ThreadPoolExecutor

    import numpy as np

    from oracle import oracle as O

    threads = os.cpu_count() or 1
    per = 1 << (3 * L)
    sample = np.random.default_rng(5).permutation(nb)
    done = 0
    t0 = time.time()
    with ThreadPoolExecutor(threads) as ex:
        while done < nb and time.time() - t0 < args.cpu_seconds * 0.5:
            chunk = sample[done:done + 16 * threads]
            list(ex.map(lambda i: ds.decode(int(i), L), chunk))
            done += len(chunk)
    port_s = time.time() - t0
    res = {"value": done * per / port_s / 1e9, "unit": "Gvoxels/s", "cores": threads, "kind": "port",
           "sample": f"oracle rANS + finest-LOD decode of {done} random bricks of the volume on {threads} threads"}
    if O.ref_available():
        R = O.ref()
        with tempfile.NamedTemporaryFile(suffix=".csbd") as f:
            f.write(data)
            f.flush()
            n = R.ref_load(f.name.encode())
        if n == nb:
            # file index of every non-constant brick (palette size > 1), random order
            pal = np.array([ds.brick(i).P for i in range(nb)])
            file_of = np.zeros(nb, np.int64)
            for k in range(nb):
                file_of[R.ref_brick_id(k)] = k
            cand = file_of[np.random.default_rng(6).permutation(np.nonzero(pal > 1)[0])].astype(np.uint32)
            dec, pos, cs, rej = 0, 0, ctypes.c_uint64(), ctypes.c_uint32()
            t1 = time.time()
            while pos < cand.size and time.time() - t1 < args.cpu_seconds * 0.5:
                part = np.ascontiguousarray(cand[pos:pos + 64 * threads])
                got = R.ref_decode_finest_batch(part.ctypes.data_as(O.c_u32p), part.size, 32, threads,
                                                ctypes.byref(cs), ctypes.byref(rej))
                dec += max(0, int(got))
                pos += part.size
            ref_s = time.time() - t1
            res["reference"] = {
                "value": dec * per / ref_s / 1e9, "unit": "Gvoxels/s", "cores": threads, "kind": "reference",
                "sample": f"reference decodeRansStream + decodeBrick<32> (oracle/_ref) on {dec} random non-constant "
                          f"bricks on {threads} threads"}
            # the port over the same kind of bricks, for the calibration
            nc = np.random.default_rng(6).permutation(np.nonzero(pal > 1)[0])
            done2 = 0
            t2 = time.time()
            with ThreadPoolExecutor(threads) as ex:
                while done2 < nc.size and time.time() - t2 < min(ref_s, args.cpu_seconds * 0.5):
                    chunk = nc[done2:done2 + 16 * threads]
                    list(ex.map(lambda i: ds.decode(int(i), L), chunk))
                    done2 += len(chunk)
            port_nc = done2 * per / (time.time() - t2) / 1e9
            # the reference's own decoder is the baseline; the port beside it
            ref_res = res.pop("reference")
            ref_res["port"] = dict(res, value_on_nonconstant_bricks=port_nc)
            return ref_res
    return res


def bench_config4(args):
    print(json.dumps(run_config4(args)))


def run_config4(args):
    """1024^3 u32 volume (seed 3, the 512^3 recipe at the same seed density),
    32^3 bricks (32,768), direct-mapped 4 GiB brick cache. Camera paths are
    Catmull-Rom splines through 8 seeded random control points inside the
    volume, 600 frames at 1920x1080 each. The plain marcher's visit logs of
    `train_paths` paths train the prefetch MLP (12 pose features + 32^3
    visited mask -> 512 -> 512 -> 32^3, FP32, on the GPU); `eval_paths`
    held-out paths are then rendered from a cold cache three ways: plain
    on-demand (fcs_render: decode the frustum's non-resident bricks, then
    march), visibility-driven on-demand (fcs_render_visible) and learned
    prefetch (the network every `every` frames, fcs_render_prefetch).
    value = 1080p frames/s of learned prefetch over the held-out paths (wall
    clock per frame, host work included); every frame's checksum is compared
    with the plain form's."""
    import numpy as np

    import fcs
    from fcs import paths
    from fcs import prefetch as PF

    W, H = 1920, 1080
    t0 = time.time()
    gen = fcs.Generator(4)
    data = fcs.encode_generated(gen, 32)
    log(f"[bench] generated + encoded {gen.size}^3 volume: {len(data)} bytes in {time.time() - t0:.1f}s")
    ctx = fcs.Context(0)
    vol = ctx.load(data)
    if np.any(vol.decode_rans()):
        raise RuntimeError("corrupt bricks")
    vol.cache_reserve_direct()
    nb, G = vol.info["num_bricks"], float(vol.info["grid_size"])
    horizon = 3 * args.every
    samples = []
    t1 = time.time()
    for p in range(args.train_paths):
        feats, vis = PF.collect_visits(vol, paths.flythrough(3, p, G, frames=args.frames), W, H)
        samples.append(PF.build_samples(feats, vis, horizon))
        log(f"[bench] visit log of training path {p}: {vis.sum(1).mean():.0f} bricks visited per frame")
    collect_s = time.time() - t1
    mlp, losses, train_s, n_samples = PF.train(ctx, nb, samples, hidden=512, epochs=args.epochs, batch=64, lr=1e-3)
    log(f"[bench] trained on {n_samples} samples in {train_s:.2f}s, loss {losses[0]:.4f} -> {losses[-1]:.4f}")
    runs = {"plain": [], "visible": [], "learned": []}
    for p in range(args.train_paths, args.train_paths + args.eval_paths):
        cams = paths.flythrough(3, p, G, frames=args.frames)
        runs["plain"].append(PF.run_path(vol, cams, W, H, mode="plain"))
        runs["visible"].append(PF.run_path(vol, cams, W, H, mode="visible"))
        runs["learned"].append(PF.run_path(vol, cams, W, H, mlp=mlp, threshold=args.threshold, max_prefetch=4096,
                                           every=args.every, ahead=args.every))
        log(f"[bench] held-out path {p} done")
    learned = PF.compare(runs["plain"], runs["learned"])
    visible = PF.compare(runs["plain"], runs["visible"])
    frames = learned["frames"]
    out = {
        "metric": "1080p frames/s, learned prefetch vs plain on-demand GPU form",
        "value": learned["fps_prefetch"],
        "unit": "frames/s",
        "n_gpus": 1, "steps": frames, "warmup": 0,
        "ms_per_step": 1000.0 / learned["fps_prefetch"],
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "u32 labels, fp32 network",
        "data": "synthetic",
        "config": {"workload": f"{gen.size}^3 u32 Voronoi (seed 3), 32^3 bricks x{nb}, direct-mapped 4 GiB cache; "
                               f"{args.train_paths} training + {args.eval_paths} held-out Catmull-Rom paths x "
                               f"{args.frames} frames at 1920x1080",
                   "network": "12 + 32768 -> 512 -> 512 -> 32768 FP32",
                   "threshold": args.threshold, "every": args.every, "horizon": horizon},
        "fps_plain_on_demand": learned["fps_plain"],
        "fps_visibility_driven": visible["fps_prefetch"],
        "fps_learned_prefetch": learned["fps_prefetch"],
        "speedup_vs_plain": learned["speedup"],
        "speedup_visibility_driven_vs_plain": visible["speedup"],
        # device time of the frame's own work (decode + march; the network's
        # forward pass and the prefetch decodes run outside it)
        "device_ms_per_frame": {"plain": learned["device_ms_per_frame_plain"],
                                "visibility_driven": visible["device_ms_per_frame_prefetch"],
                                "learned": learned["device_ms_per_frame_prefetch"]},
        "miss_rate": learned["miss_rate"],
        "miss_rate_visibility_driven": visible["miss_rate"],
        "on_demand_decodes": {"plain": learned["on_demand_plain"], "visibility_driven": visible["on_demand_prefetch"],
                              "learned": learned["on_demand_prefetch"]},
        "prefetched": learned["prefetched"], "wasted_prefetches": learned["wasted"],
        "wasted_frac": learned["wasted_frac"],
        "label_mismatch_frames": learned["label_mismatch_frames"] + visible["label_mismatch_frames"],
        "train": {"samples": n_samples, "epochs": args.epochs, "seconds": train_s, "loss_first": float(losses[0]),
                  "loss_last": float(losses[-1]), "visit_log_seconds": collect_s},
    }
    mlp.close()
    vol.close()
    ctx.close()
    return out

def bench_config5(args):
    """Serving: 64 viewer sessions over four volumes of the 512^3 recipe at
    the same seed density (256^3, 512^3, 1024^3, 2048^3 u32, 32^3 bricks),
    admitted under a 200 GB HBM budget (all four fit whole: every brick at
    every LOD resident, 41.8 GiB). Requests arrive as one Poisson stream of
    2,000 frames/s for `duration` seconds, session uniform, session s on
    volume s % 4 with a random-walk orbit camera, resolution uniform over
    640x480 / 1280x720 / 1920x1080. Each request is a frame whose first-hit
    label image is copied to pinned host memory; latency = time the host sees
    it - arrival. Served on `views` render targets (HIP streams) with
    hipGraph-replayed frames; the same trace is also served with eagerly
    launched frames and with shaded RGBA frames (transfer function, shading,
    shadow rays; the image copied back is RGBA8), and the whole trace released
    at once measures capacity.
    value = frames/s served at the offered rate (first-hit frames, graph form)."""
    import numpy as np

    import fcs
    from fcs import serve

    sizes = [int(x) for x in args.sizes.split(",") if x]
    ctx = fcs.Context(0)
    vols = []
    for s in sizes:
        t0 = time.time()
        gen = fcs.Generator(5, size=s)
        data = fcs.encode_generated(gen, 32, progress=s >= 1024)
        vols.append(ctx.load(data))
        log(f"[bench] serving volume {s}^3: {len(data)} bytes, generated + encoded + loaded in {time.time() - t0:.1f}s")
        del data
    t0 = time.time()
    plan, pool = serve.admit(vols, [1.0] * len(vols), int(args.budget_gb * 1e9))
    ctx.sync()
    admit_s = time.time() - t0
    log(f"[bench] admission plan {[int(p) for p in plan]} in {admit_s:.1f}s")

    def make_views(graph, rgba=False):
        return [serve.View(ctx, 1920, 1080, graph=graph, labels=True, checksum=False, rgba=rgba)
                for _ in range(args.views)]

    def warm(views):
        # one frame per (view, volume, resolution): graphs captured, code loaded
        for v in views:
            for vol in vols:
                cam = vol.orbit_camera(0.5, 0.3, 1.2 * math.sqrt(3) * vol.info["grid_size"])
                for w, h in serve.RESOLUTIONS:
                    v.render_async(vol, cam, w, h)
                    v.wait()

    reqs, meta = serve.make_trace(vols, n_sessions=args.sessions, rate=args.rate, duration=args.duration, seed=17)
    n = len(meta["arrival_s"])
    log(f"[bench] trace: {n} requests over {args.duration}s from {args.sessions} sessions")
    out = {}
    import ctypes
    burst = (type(reqs[0]) * n)()
    ctypes.memmove(burst, reqs, ctypes.sizeof(burst))
    for i in range(n):
        burst[i].arrival_s = 0.0
    px = float((meta["w"] * meta["h"]).sum())
    # first-hit label frames with graph replay / eager launches, then shaded
    # RGBA frames (graph replay); capacity = the same requests released at once
    for name, graph, rgba in (("graph", True, False), ("eager", False, False), ("shaded", True, True)):
        views = make_views(graph, rgba)
        warm(views)
        res = serve.run(views, vols, reqs)
        st = serve.latency_stats(meta, res)
        st["graphs_per_view"] = views[0].graphs
        out[name] = st
        log(f"[bench] {name}: p50 {st['p50_ms']:.3f} ms p99 {st['p99_ms']:.3f} ms {st['frames_per_s']:.0f} frames/s")
        if graph:
            t1 = time.perf_counter()
            rb = serve.run(views, vols, burst)
            wall = time.perf_counter() - t1
            out["capacity" if name == "graph" else "shaded_capacity"] = {
                "frames_per_s": n / wall, "rays_per_s": px / wall,
                "frames_1080p_equiv_per_s": px / wall / (1920 * 1080),
                "device_ms_mean": float(rb["device_ms"].mean())}
            log(f"[bench] {name} capacity: {n / wall:.0f} frames/s, {px / wall / 1e9:.2f} Grays/s")
        for v in views:
            v.close()
    g = out["graph"]
    line = {
        "metric": "served frames/s at 2000 offered, p50/p99 latency",
        "value": g["frames_per_s"],
        "unit": "frames/s",
        "n_gpus": 1, "steps": n, "warmup": args.views * len(vols) * 3,
        "ms_per_step": 1000.0 / g["frames_per_s"] if g["frames_per_s"] else None,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "u32", "data": "synthetic",
        "config": {"workload": f"serving {args.sessions} sessions over {'/'.join(str(s) + '^3' for s in sizes)} u32 "
                               f"volumes (512^3 recipe), 32^3 bricks, Poisson {args.rate:.0f} frames/s for "
                               f"{args.duration}s, resolutions 640x480/1280x720/1920x1080, {args.views} views",
                   "budget_gb": args.budget_gb, "plan": [int(p) for p in plan],
                   "cache_gib": float(sum(int(x) for x in pool)) / 2**30},
        "p50_ms": g["p50_ms"], "p99_ms": g["p99_ms"],
        "graph": out["graph"], "eager": out["eager"], "capacity": out["capacity"],
        "shaded": out["shaded"], "shaded_capacity": out["shaded_capacity"],
        "admission_s": admit_s,
        "roofline": None, "cpu_baseline": None,
    }
    print(json.dumps(line))


if __name__ == "__main__":
    main()
