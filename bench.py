"""Benchmark: primary visibility at 1920x1080 over 512 noise-filled objects.

    python bench.py [--gpus 1] [--steps K] [--warmup W]

One step = one frame of the hot path on one GPU: cull/bin of every cluster
pointer, tile lists, and the per-tile 8^3 DDA producing the bit-exact 64-bit
visibility buffer (tg's visibility.frag). Inputs (scene arrays) are resident
in HBM before the timed region; the output stays in HBM.

Prints one JSON line with rays/s (= primary rays = pixels), the roofline of
the dominant kernel (vis_traverse, HIP events on the library's stream) and
the CPU oracle timed on this host as the baseline.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time
from pathlib import Path

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

import numpy as np  # noqa: E402

import tg_amd as T  # noqa: E402

W, H = 1920, 1080
HBM_PEAK_GBS = 8000.0
DOMINANT = "vis_traverse"


def algorithmic_bytes(stats: T.VisibilityStats, n_objects: int, n_pixels: int) -> int:
    """Per frame: 8 B written per pixel; per cluster that survives the cull
    its 64 B of solid bits plus its 4 B pointer-map and 4 B cluster->object
    entries; 96 B per object record."""
    return 8 * n_pixels + 72 * int(stats.clusters_binned) + 96 * n_objects


def cpu_baseline(scene: T.Scene, params, budget_s: float = 12.0) -> dict:
    import oracle  # the CPU baseline leg is the only bench code that uses the oracle

    oscene = oracle.OracleScene(scene)
    threads = int(os.environ.get("OMP_NUM_THREADS", "0")) or (os.cpu_count() or 1)
    threads = min(threads, 16)
    res = {}
    for label, n_thr, rows in (("all", threads, 1), ("one", 1, 8)):
        t0 = time.perf_counter()
        rays = 0
        ci = 0
        while True:
            cam = T.generate_camera(params, ci, W, H)
            oracle.visibility_culled(oscene, cam, W, H, row_step=rows, n_threads=n_thr)
            rays += W * ((H + rows - 1) // rows)
            ci += 1
            if time.perf_counter() - t0 > budget_s / 2:
                break
        res[label] = (rays / (time.perf_counter() - t0), ci, rows)
    all_rate, n_frames, _ = res["all"]
    one_rate, n1, rows1 = res["one"]
    return {
        "value": all_rate, "unit": "rays/s", "cores": threads, "kind": "port",
        "sample": f"oracle culled visibility, {n_frames} full 1920x1080 frames (seeded cameras) on {threads} threads; "
                  f"single-core rate {one_rate:.4g} rays/s over every {rows1}th row of {n1} frames",
        "single_core_value": one_rate,
    }


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--seed", type=int, default=1)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU: this build has no multi-GPU path (each GPU would render whole "
              "cameras or screen tiles of its own; see the sharding design). Run with --gpus 1.", file=sys.stderr)
        return 2

    params = T.noise_objects_params(seed=args.seed)
    scene = T.generate_scene(params)
    cam = T.generate_camera(params, 0, W, H)
    ctx = T.Context(0)
    ctx.upload(scene)
    out = ctx.alloc(W * H * 8)

    for _ in range(args.warmup):
        ctx.visibility_device(cam, W, H, out)
    ctx.synchronize()

    ctx.profile(True)
    ctx.profile_reset()
    ctx.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        ctx.visibility_device(cam, W, H, out)
    ctx.synchronize()
    elapsed = time.perf_counter() - t0
    prof = ctx.profile_stats()
    ctx.profile(False)
    stats = ctx.last_stats()

    ms_per_step = 1e3 * elapsed / args.steps
    rays_per_s = W * H * args.steps / elapsed
    launches, total_ms = prof[DOMINANT]
    avg_ms = total_ms / launches
    b_alg = algorithmic_bytes(stats, scene.objects.size, W * H)
    achieved = b_alg / (avg_ms * 1e-3) / 1e9

    line = {
        "metric": "rays/s (primary DDA + SVO bounce, 1080p); neural vs plain GI at <=2% rel-MSE",
        "value": rays_per_s,
        "unit": "rays/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {
            "workload": "primary visibility, 512 noise-filled objects of 1-8 clusters per axis in a 512^3 world, "
                        "1920x1080, camera at the world centre (seeded yaw/pitch), bit-exact 64-bit visibility",
            "objects": int(scene.objects.size), "clusters": scene.n_clusters, "width": W, "height": H,
            "cameras": 1, "seed": args.seed, "rays_counted": "primary (one per pixel)",
        },
        "roofline": {
            "bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS, "traffic": None,
            "kernel": DOMINANT, "kernel_avg_ms": avg_ms, "alg_bytes_per_launch": b_alg,
        },
        "kernels_ms_per_step": {k: v[1] / args.steps for k, v in prof.items()},
        "cull": {"clusters_binned": int(stats.clusters_binned), "tile_entries": int(stats.tile_entries)},
    }
    ctx.free(out)
    ctx.close()
    if not args.no_cpu_baseline:
        line["cpu_baseline"] = cpu_baseline(scene, params)
    print(json.dumps(line), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
