"""Benchmark: primary visibility + one-bounce SVO GI at 1920x1080 over 512
noise-filled objects (the metric's workload).

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload frame|visibility|batch]

One step = one frame of the hot path on one GPU:
  - primary visibility: cull/bin of every cluster pointer, tile lists and
    the per-tile 8^3 DDA producing the bit-exact 64-bit visibility buffer
    (tg's visibility.frag);
  - one-bounce GI through the depth-11 SVO built around the camera: 4
    cosine-weighted bounce rays per visible pixel and one shadow ray per
    sun-facing bounce hit.
The scene arrays and the SVO are resident in HBM before the timed region
(the SVO build is timed separately); outputs stay in HBM.

Prints one JSON line: rays/s (primary + bounce + shadow rays per frame over
the frame time), the roofline of the dominant kernel (HIP events on the
library's stream) and the CPU oracle timed on this host as the baseline.

The same line carries the neural-cache frame on the same view (`neural`)
and the batched stress frame (`batch`): 4096 objects of 1-32 clusters per
axis in a 4096^3 world, a depth-12 SVO, 16 cameras at 1080p, each form
captured as one hipGraph after autotuning, with the oracle's visibility of
sampled rows of every camera compared word for word.
"""
from __future__ import annotations

import argparse
import ctypes
import json
import math
import os
import shutil
import signal
import sqlite3
import subprocess
import sys
import tempfile
import time
from pathlib import Path

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

import numpy as np  # noqa: E402

import tg_amd as T  # noqa: E402

W, H = 1920, 1080
SPP = 4
DEPTH = 11
HBM_PEAK_GBS = 8000.0


def vis_alg_bytes(stats, n_objects: int, n_pixels: int) -> int:
    """Visibility frame: 8 B written per pixel; per cluster surviving the
    cull its 64 B of solid bits plus 4 B pointer-map and 4 B cluster->object
    entries; 96 B per object record."""
    return 8 * n_pixels + 72 * int(stats.clusters_binned) + 96 * n_objects


def trace_alg_bytes(gst, info) -> int:
    """The frame's SVO trace launch (bounce rays with their shadow rays
    traced in the same launch): per bounce ray its 32-B queue record read;
    per shadow ray the 2-B material of the bounce hit read and the 12-B
    contribution written; the SVO's nodes (32 B each) and leaf bit blocks
    (4 KiB each) read once."""
    return (32 * int(gst["bounce_rays"]) + 14 * int(gst["shadow_rays"])
            + 32 * int(info["n_nodes"]) + 4096 * int(info["n_leaves"]))


def svo_memory(ctx, info, out: dict) -> dict:
    """Device bytes of the SVO's data, beside what a dense per-voxel material
    grid (2 B for each of a leaf's 32^3 voxels) would add instead of the
    solid voxels' materials and the word prefixes."""
    dev, solid = ctx.svo_device_bytes()
    nl = int(info["n_leaves"])
    dense = dev - 2 * solid - 2 * 1024 * nl + 2 * 32768 * nl
    out.update({"device_bytes": dev, "solid_voxels": solid, "device_bytes_dense_materials": dense,
                "device_fraction_of_dense": dev / dense if dense else None})
    return out


MFMA_BF16_PEAK_TFLOPS = 2500.0
# Cache fallback: trace a pixel when its cell's squared bias exceeds this
# multiple of the plain estimator's variance (a margin for the noise of the
# cells' estimates).
NRC_FALLBACK_THRESHOLD = 16.0   # tracer rule: trace where bias^2 > 16 x the 4-spp variance
NRC_FALLBACK_MIN_SAMPLES = 0    # no fallback for cells merely short of samples
NRC_TRAIN_BATCH = 1 << 14   # samples per Adam step: 4 steps per frame of 65536 training samples
NRC_EMA_DECAY = 0.996       # inference weights: moving average over ~250 Adam steps (~60 frames)
NRC_SIDE_BATCH = 1 << 12    # the side form with 16 steps of 4096 samples per frame
NRC_SIDE_EMA_DECAY = 0.999  # its moving average over the same ~60 frames
NRC_REF_FRAMES = 4096       # 4-spp frames in the error reference (16384 spp)
NRC_SIDE_FRAMES = 3200      # training frames of the forms reported beside the bench's own
NRC_PATH_DEG_PER_FRAME = 0.05   # moving camera: yaw turned per frame
NRC_PATH_MARKS = (200, 800, 3200)
NRC_PATH_REF_FRAMES = 256   # 1024-spp reference at each mark of the path
NRC_FWD_FLOPS = 2 * (78 * 64 + 3 * 64 * 64 + 64 * 3)   # per sample, unpadded
NRC_TRAIN_FLOPS = 3 * NRC_FWD_FLOPS                     # forward + two backward products


def rel_mse(x, ref) -> float:
    x = np.asarray(x, np.float64)
    ref = np.asarray(ref, np.float64)
    return float(((x - ref) ** 2).sum() / (ref ** 2).sum())


def moving_camera(ctx, sun, vis_d, rad_d, irr_d) -> dict:
    """The bench's cache form trained online along a camera path: from the
    world centre, pitch -0.15 rad, the yaw turning NRC_PATH_DEG_PER_FRAME
    per frame. At each mark, the last frame's irradiance against a 1024-spp
    reference at that frame's pose (two 512-spp halves, debiased as above)."""
    step = math.radians(NRC_PATH_DEG_PER_FRAME)

    def cam_at(k: int) -> np.ndarray:
        return T.make_camera((256.0, 256.0, 256.0), -0.15, 0.6 + k * step, 0.0, math.radians(60.0), W / H,
                             0.1, 1000.0)

    ctx.nrc_create(T.nrc_config(origin=(0.0, 0.0, 0.0), scale=1.0 / 512, train_samples=1 << 16,
                                train_batch=NRC_TRAIN_BATCH, train_spp=4, fallback_threshold=NRC_FALLBACK_THRESHOLD,
                                fallback_min_samples=NRC_FALLBACK_MIN_SAMPLES, cell_size=8,
                                flags=T.NRC_FLAG_FALLBACK_VS_TRACER, ema_decay=NRC_EMA_DECAY))
    prm = T.gi_params(sun, spp=SPP, max_distance=1024.0, frame=0)
    log(f"neural: moving camera, {max(NRC_PATH_MARKS)} frames at {NRC_PATH_DEG_PER_FRAME} deg per frame")
    marks, k = [], 0
    for mark in NRC_PATH_MARKS:
        while k < mark:
            ctx.nrc_frame_device(cam_at(k), W, H, prm, vis_d, rad_d, irr_d)
            k += 1
        cam = cam_at(k - 1)
        irr = np.empty((H, W, 3), np.float32)
        vis = np.empty((H, W), np.uint64)
        ctx.d2h(irr, irr_d)
        ctx.d2h(vis, vis_d)
        fb = int(ctx.nrc_stats()["fallback_pixels"])
        halves = []
        for base in (300000, 400000):
            acc = np.zeros((H, W, 3))
            for f in range(NRC_PATH_REF_FRAMES // 2):
                ctx.frame_device(cam, W, H, T.gi_params(sun, spp=SPP, max_distance=1024.0, frame=base + f),
                                 vis_d, rad_d, irr_d)
                out = np.empty((H, W, 3), np.float32)
                ctx.d2h(out, irr_d)
                acc += out
            halves.append(acc / (NRC_PATH_REF_FRAMES // 2))
        m = vis != T.VISIBILITY_EMPTY
        ref = 0.5 * (halves[0] + halves[1])
        ref_var = rel_mse(halves[0][m], halves[1][m]) / 4.0
        e = rel_mse(irr[m], ref[m])
        marks.append({"frames": mark, "yaw_turned_deg": (mark - 1) * NRC_PATH_DEG_PER_FRAME, "rel_mse": e,
                      "rel_mse_debiased": e - ref_var, "reference_rel_variance": ref_var, "fallback_pixels": fb})
    return {"deg_per_frame": NRC_PATH_DEG_PER_FRAME, "reference_spp": 4 * NRC_PATH_REF_FRAMES, "marks": marks}


def neural_bench(ctx, cam, sun, vis_d, rad_d, irr_d, steps: int, pretrain: int, plain_ms: float,
                 plain_rays: int) -> dict:
    """The neural irradiance cache (512-object noise scene, 1920x1080) on the
    same view, trained online for `pretrain` frames of a static camera.
    Errors are relative MSE against a 16384-spp FP64-accumulated reference
    (4096 independent 4-spp frames of the plain tracer) in two independent
    8192-spp halves: a quarter of the halves' rel-MSE is the reference's own
    variance, which every measured (raw) error carries and the debiased
    figures subtract (the cache is trained on other samples, so its error
    and the reference's noise add). The 2 % criterion is taken on the raw
    error. A moving camera (the yaw turning NRC_PATH_DEG_PER_FRAME per
    frame) gives the error a cache reaches while its view changes."""
    def irr_frame(frame: int) -> np.ndarray:
        ctx.frame_device(cam, W, H, T.gi_params(sun, spp=SPP, max_distance=1024.0, frame=frame), vis_d, rad_d, irr_d)
        out = np.empty((H, W, 3), np.float32)
        ctx.d2h(out, irr_d)
        return out

    log(f"neural: {NRC_REF_FRAMES}-frame reference")
    halves = []
    for base in (100000, 200000):
        acc = np.zeros((H, W, 3))
        for f in range(NRC_REF_FRAMES // 2):
            acc += irr_frame(base + f)
        halves.append(acc / (NRC_REF_FRAMES // 2))
    vis = np.empty((H, W), np.uint64)
    ctx.d2h(vis, vis_d)
    mask = vis != T.VISIBILITY_EMPTY
    ref = 0.5 * (halves[0] + halves[1])
    ref_var = rel_mse(halves[0][mask], halves[1][mask]) / 4.0

    def err(x: np.ndarray) -> dict:
        m = rel_mse(x[mask], ref[mask])
        return {"rel_mse": m, "rel_mse_debiased": m - ref_var}

    plain = irr_frame(7)
    prm = T.gi_params(sun, spp=SPP, max_distance=1024.0, frame=0)

    def run(flags: int, threshold: float, min_samples: int, ema: float, milestones=(), frames: int = pretrain,
            batch: int = NRC_TRAIN_BATCH):
        """a fresh cache trained for `frames` untimed frames (errors read at
        the milestones), then `steps` timed frames"""
        ctx.nrc_create(T.nrc_config(origin=(0.0, 0.0, 0.0), scale=1.0 / 512, train_samples=1 << 16,
                                    train_batch=batch, train_spp=4, fallback_threshold=threshold,
                                    fallback_min_samples=min_samples, cell_size=8, flags=flags, ema_decay=ema))
        log(f"neural: flags {flags} threshold {threshold} ema {ema}: {frames} training frames")
        curve = []
        for k in range(1, frames + 1):
            ctx.nrc_frame_device(cam, W, H, prm, vis_d, rad_d, irr_d)
            if k in milestones:
                out = np.empty((H, W, 3), np.float32)
                ctx.d2h(out, irr_d)
                curve.append({"frames": k, **err(out), "fallback_pixels": int(ctx.nrc_stats()["fallback_pixels"])})
        ctx.synchronize()
        t0 = time.perf_counter()
        for _ in range(steps):
            ctx.nrc_frame_device(cam, W, H, prm, vis_d, rad_d, irr_d)
        ctx.synchronize()
        elapsed = time.perf_counter() - t0
        # per-kernel times from event-profiled frames after the timed ones
        ctx.profile(True)
        ctx.profile_reset()
        for _ in range(steps):
            ctx.nrc_frame_device(cam, W, H, prm, vis_d, rad_d, irr_d)
        ctx.synchronize()
        prof = ctx.profile_stats()
        ctx.profile(False)
        st = ctx.nrc_stats()
        irr = np.empty((H, W, 3), np.float32)
        ctx.d2h(irr, irr_d)
        return 1e3 * elapsed / steps, prof, st, irr, curve

    # reported beside: the t^2 rule, no fallback, no weight average; then the
    # bench's form, whose error is also read while it trains
    side = min(pretrain, NRC_SIDE_FRAMES)
    ms_t2, _, st_t2, irr_t2, _ = run(0, 1.0, 1, NRC_EMA_DECAY, frames=side)
    ms_nf, _, _, irr_nf, _ = run(T.NRC_FLAG_NO_FALLBACK, 1.0, 1, NRC_EMA_DECAY, frames=side)
    ms_ne, _, st_ne, irr_ne, _ = run(T.NRC_FLAG_FALLBACK_VS_TRACER, NRC_FALLBACK_THRESHOLD, NRC_FALLBACK_MIN_SAMPLES,
                                     0.0, frames=side)
    # the bench's form with 16 Adam steps of 4096 samples per frame instead of
    # 4 of 16384 (the moving average over the same number of frames), at the
    # side forms' frame count (its error against the bench form's
    # convergence entry at the same count)
    ms_b4, _, st_b4, irr_b4, _ = run(T.NRC_FLAG_FALLBACK_VS_TRACER, NRC_FALLBACK_THRESHOLD,
                                     NRC_FALLBACK_MIN_SAMPLES, NRC_SIDE_EMA_DECAY, frames=side, batch=NRC_SIDE_BATCH)
    marks = sorted({m for m in (200, 400, 800, 1600, 3200, 6400, 12800) if m < pretrain} | {pretrain})
    ms, prof, st, irr, curve = run(T.NRC_FLAG_FALLBACK_VS_TRACER, NRC_FALLBACK_THRESHOLD, NRC_FALLBACK_MIN_SAMPLES,
                                   NRC_EMA_DECAY, marks)
    e_neural, e_plain = err(irr), err(plain)
    # the plain estimator's error falls as 1/spp: spp and frame time it
    # needs for the cache's error (trace time taken as proportional to spp)
    spp_eq = SPP * e_plain["rel_mse_debiased"] / e_neural["rel_mse_debiased"] if e_neural["rel_mse_debiased"] > 0 else None
    def graph_ms(flags: int) -> float:
        """one-camera frame captured as a graph (the cache keeps training on
        every replay): 3 untimed replays, then `steps` timed ones"""
        g = ctx.frame_graph(cam, W, H, prm, vis_d, rad_d, irr_d, flags)
        try:
            for _ in range(3):
                g.launch()
            ctx.synchronize()
            t0 = time.perf_counter()
            for _ in range(steps):
                g.launch()
            ctx.synchronize()
            return 1e3 * (time.perf_counter() - t0) / steps
        finally:
            g.destroy()

    captured = {"plain_frame_ms": graph_ms(0), "neural_frame_ms": graph_ms(T.FRAME_NEURAL)}
    moving = moving_camera(ctx, sun, vis_d, rad_d, irr_d)
    captured["speedup_vs_plain"] = captured["plain_frame_ms"] / captured["neural_frame_ms"]
    k_ms = {k: v[1] / steps for k, v in prof.items()}
    visible = int(st["visible_pixels"])
    infer_ms = k_ms.get("nrc_infer_k", 0.0)
    train_ms = sum(k_ms.get(k, 0.0) for k in ("nrc_fwdbwd_k", "nrc_wgrad_k", "nrc_adam_k"))
    infer_tflops = NRC_FWD_FLOPS * visible / (infer_ms * 1e-3) / 1e12 if infer_ms else None
    train_tflops = NRC_TRAIN_FLOPS * int(st["train_samples"]) / (train_ms * 1e-3) / 1e12 if train_ms else None
    traced = W * H + int(st["train_rays"]) + int(st["fallback_rays"])
    # HBM roofline of the frame's trace launch, the same algorithmic bytes as
    # the plain frame's (trace_alg_bytes): bounce rays are the listed entries
    # times spp, the rest of each count are shadow rays
    t_bounce = int(st["train_samples"]) * 4  # train_spp of the cache configuration above
    f_bounce = int(st["fallback_pixels"]) * SPP
    rays_b = t_bounce + f_bounce
    rays_s = int(st["train_rays"]) + int(st["fallback_rays"]) - rays_b
    info = ctx.svo_info()
    n_launch, tr_ms = prof.get("gi_trace_k", (0, 0.0))
    tr_avg = tr_ms / n_launch if n_launch else None
    alg = ((32 * rays_b + 14 * rays_s) * steps / n_launch + 32 * int(info["n_nodes"]) + 4096 * int(info["n_leaves"])
           if n_launch else None)
    achieved = alg / (tr_avg * 1e-3) / 1e9 if alg and tr_avg else None
    hbm = {"bound": "hbm", "kernel": "gi_trace_k", "kernel_avg_ms": tr_avg, "launches_per_frame": n_launch / steps,
           "alg_bytes_per_launch": alg, "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
           "frac": achieved / HBM_PEAK_GBS if achieved else None}
    return {
        "frame_ms": ms,
        "plain_frame_ms": plain_ms,
        "speedup_vs_plain": plain_ms / ms,
        "rays_traced_per_frame": traced,
        "rays_per_s_equivalent": plain_rays / (ms * 1e-3),
        "reference": {"spp": 4 * NRC_REF_FRAMES, "accumulation": "fp64", "rel_variance": ref_var},
        "rel_mse_vs_reference": {"neural": e_neural, "plain_4spp": e_plain},
        "meets_2pct_rel_mse": e_neural["rel_mse"] <= 0.02,
        "meets_2pct_rel_mse_debiased": e_neural["rel_mse_debiased"] <= 0.02,
        "plain_at_equal_error": {"spp": spp_eq, "frame_ms_est": plain_ms * spp_eq / SPP if spp_eq else None},
        "convergence": curve,
        "visible_pixels": visible, "fallback_pixels": int(st["fallback_pixels"]),
        "fallback_rule": (f"cell bias^2 > {NRC_FALLBACK_THRESHOLD} x the plain estimator's variance at 4 spp, "
                          f"min samples {NRC_FALLBACK_MIN_SAMPLES}"),
        "weight_average": f"inference on the moving average of the weights, decay {NRC_EMA_DECAY} per Adam step",
        "fallback_rule_t2": {"rule": "cell bias^2 > 1.0 x t^2, min samples 1", "frame_ms": ms_t2,
                             "fallback_pixels": int(st_t2["fallback_pixels"]), **err(irr_t2)},
        "no_fallback": {"frame_ms": ms_nf, **err(irr_nf)},
        "no_weight_average": {"frame_ms": ms_ne, "fallback_pixels": int(st_ne["fallback_pixels"]), **err(irr_ne)},
        "side_forms_frames": side,
        "train_batch_4096": {"frame_ms": ms_b4, "fallback_pixels": int(st_b4["fallback_pixels"]),
                             "adam_steps_per_frame": (1 << 16) // NRC_SIDE_BATCH, "ema_decay": NRC_SIDE_EMA_DECAY,
                             **err(irr_b4)},
        "moving_camera": moving,
        "train_samples": int(st["train_samples"]), "adam_steps_total": int(st["adam_steps"]),
        "train_loss": float(st["train_loss"]), "pretrain_frames": pretrain,
        "kernels_ms_per_step": {k: v for k, v in k_ms.items() if k.startswith("nrc_")},
        "mfma": {"infer_tflops": infer_tflops, "train_tflops": train_tflops, "peak_tflops": MFMA_BF16_PEAK_TFLOPS,
                 "infer_frac": infer_tflops / MFMA_BF16_PEAK_TFLOPS if infer_tflops else None,
                 "train_frac": train_tflops / MFMA_BF16_PEAK_TFLOPS if train_tflops else None,
                 "flops_per_sample": {"forward": NRC_FWD_FLOPS, "train": NRC_TRAIN_FLOPS}},
        "captured_graphs": captured,
        "roofline": hbm,
        "config": {"network": "96(79 used)->64x4->3 ReLU, bf16 MFMA, fp32 Adam lr 1e-3", "train_samples": 1 << 16,
                   "train_batch": NRC_TRAIN_BATCH, "train_spp": 4, "encoding": "12-frequency sin/cos of p/512",
                   "ema_decay": NRC_EMA_DECAY},
    }


# Batched stress frame: 4096 noise-filled objects of 1..32 clusters per axis
# in a 4096^3 world, 64 LUTs, a depth-12 SVO over the world, 16 seeded
# cameras at 1920x1080 in one captured frame.
BATCH_OBJECTS, BATCH_EXTENT, BATCH_WORLD, BATCH_LUTS = 4096, 32, 4096.0, 64
BATCH_DEPTH, BATCH_CAMS = 12, 16
BATCH_WORKLOAD = ("batched frame: 4096 noise-filled objects of 1-32 clusters per axis in a 4096^3 world, 64 LUTs, "
                  "depth-12 SVO over the world, 16 seeded cameras at 1920x1080 at the world centre, one captured "
                  "hipGraph per form (plain SVO GI at 4 spp; neural cache with fallback, trained jointly on "
                  "every camera's 2^16 targets per frame), autotuned launch shapes")


def batch_bench(args, steps: int, pretrain: int) -> dict:
    """Both forms of the batched frame as captured graphs: plain SVO GI and
    the neural cache with fallback. Each is autotuned (uncaptured frames),
    captured, replayed untimed (3 times for plain GI, `pretrain` times for
    the cache, which trains meanwhile), then `steps` times timed.
    Per-kernel times come from one uncaptured, event-profiled frame after
    the timed replays (same kernels and inputs)."""
    params = T.noise_objects_params(seed=args.seed, n_objects=BATCH_OBJECTS, max_extent=BATCH_EXTENT,
                                    world=BATCH_WORLD, n_luts=BATCH_LUTS)
    ctx = T.Context(0)
    log("batch: device scene generation")
    t0 = time.perf_counter()
    ctx.generate_scene_device(params)
    ctx.synchronize()
    gen_ms = 1e3 * (time.perf_counter() - t0)
    cams = np.concatenate([T.generate_camera(params, i, W, H) for i in range(BATCH_CAMS)])
    mn = np.floor(cams[0]["camera"][:3]) - float(1 << (BATCH_DEPTH - 1))
    log(f"batch: SVO build, depth {BATCH_DEPTH}")
    t0 = time.perf_counter()
    ctx.svo_build(mn, BATCH_DEPTH)
    build_ms = 1e3 * (time.perf_counter() - t0)
    info = ctx.svo_info()
    prm = T.gi_params(T.seeded_sun(args.seed), spp=SPP, max_distance=1024.0)
    n_px = BATCH_CAMS * W * H
    vis_d, rad_d, irr_d = ctx.alloc(8 * n_px), ctx.alloc(12 * n_px), ctx.alloc(12 * n_px)
    sizes = T._GenSizes()
    T.load_library().tgx_gen_scene_sizes(ctypes.byref(params), ctypes.byref(sizes))
    out = {"scene": {"objects": BATCH_OBJECTS, "clusters": int(sizes.n_clusters), "world": BATCH_WORLD,
                     "luts": BATCH_LUTS, "generate_ms": gen_ms},
           "svo": svo_memory(ctx, info, {"depth": BATCH_DEPTH, "build_ms": build_ms, "nodes": int(info["n_nodes"]),
                                         "leaves": int(info["n_leaves"])}),
           "cameras": BATCH_CAMS, "width": W, "height": H, "spp": SPP}
    plain_rays = None
    for mode, flags in (("plain", 0), ("neural", T.FRAME_NEURAL)):
        if flags:
            # 16 Adam steps per frame: the batch trains for only `pretrain`
            # frames before it is timed
            ctx.nrc_create(T.nrc_config(origin=(0.0, 0.0, 0.0), scale=1.0 / BATCH_WORLD, train_samples=1 << 16,
                                        train_batch=NRC_SIDE_BATCH, train_spp=4, fallback_threshold=NRC_FALLBACK_THRESHOLD,
                                        fallback_min_samples=NRC_FALLBACK_MIN_SAMPLES, cell_size=8,
                                        flags=T.NRC_FLAG_FALLBACK_VS_TRACER | T.NRC_FLAG_JOINT_TRAINING,
                                        ema_decay=NRC_SIDE_EMA_DECAY))
        if not flags:
            # the cache learns while candidates run, so their times are not
            # comparable: the cache form keeps the plain form's tuning (the
            # same visibility and trace kernels)
            log(f"batch {mode}: autotune")
            t0 = time.perf_counter()
            best = ctx.autotune(cams, W, H, prm, vis_d, rad_d, irr_d, flags=flags, frames=1)
            tune_s = time.perf_counter() - t0
        else:
            tune_s = 0.0
        untimed = pretrain if flags else 3  # the cache trains while it replays; plain GI only warms up
        log(f"batch {mode}: capture, {untimed} untimed replays")
        t0 = time.perf_counter()
        g = None if args.batch_no_graph else ctx.frame_graph(cams, W, H, prm, vis_d, rad_d, irr_d, flags=flags)
        capture_s = time.perf_counter() - t0

        def run():
            if g is None:
                ctx.frame_batch(cams, W, H, prm, vis_d, rad_d, irr_d, flags=flags)
            else:
                g.launch()
        try:
            for _ in range(untimed):
                run()
            ctx.synchronize()
            log(f"batch {mode}: {steps} timed {'frames' if g is None else 'replays'}")
            t0 = time.perf_counter()
            for _ in range(steps):
                run()
            ctx.synchronize()
            elapsed = time.perf_counter() - t0
            st = (ctx.frame_batch(cams, W, H, prm, vis_d, rad_d, irr_d, flags=flags, stats=True) if g is None
                  else g.stats())
            if flags == 0:
                vis_host = np.empty((BATCH_CAMS, H, W), np.uint64)
                ctx.d2h(vis_host, vis_d)
        finally:
            if g is not None:
                g.destroy()
        ctx.profile(True)
        ctx.profile_reset()
        ctx.frame_batch(cams, W, H, prm, vis_d, rad_d, irr_d, flags=flags)
        ctx.synchronize()
        prof = ctx.profile_stats()
        ctx.profile(False)
        ms = 1e3 * elapsed / steps
        traced_gi = int(st["bounce_rays"]) + int(st["shadow_rays"]) + int(st["train_rays"])
        traced = int(st["primary_rays"]) + traced_gi
        if flags == 0:
            plain_rays = traced
        k_ms = {k: v[1] for k, v in prof.items()}
        dominant = max(k_ms, key=k_ms.get)
        launches, total_ms = prof[dominant]
        avg_ms = total_ms / launches
        if dominant == "gi_trace_k":
            # per bounce ray its 32-B queue record read; per shadow ray (and
            # per training ray, bounce or shadow) at least a 2-B material read
            # and a 12-B contribution write; the SVO reads (at most 8 B per
            # inner node + 4104 B per leaf per launch) are left out: a lower
            # bound
            alg = (32 * int(st["bounce_rays"]) + 14 * (int(st["shadow_rays"]) + int(st["train_rays"]))) / launches
        else:
            alg = None
        achieved = alg / (avg_ms * 1e-3) / 1e9 if alg else None
        out[mode] = {
            "frame_ms": ms,
            "rays_traced_per_frame": traced,
            "rays_traced_per_s": traced / (ms * 1e-3),
            "stats": {k: int(st[k]) for k in st.dtype.names},
            "tuning": {k: int(best[k][0]) for k in best.dtype.names},
            "tuning_source": "the plain form's autotune" if flags else "autotune of this form",
            "autotune_s": tune_s, "capture_s": capture_s, "captured": g is not None, "untimed_replays": untimed,
            "kernels_ms_per_frame": k_ms,
            "roofline": {"bound": "hbm", "kernel": dominant, "kernel_avg_ms": avg_ms, "launches_per_frame": launches,
                         "alg_bytes_per_launch": alg, "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                         "frac": achieved / HBM_PEAK_GBS if achieved else None},
        }
        if flags:
            # joint training: every camera's 2^16 targets per frame
            n_train = len(cams) * (1 << 16)
            t_ms = sum(k_ms.get(k, 0.0) for k in ("nrc_fwdbwd_k", "nrc_wgrad_k", "nrc_adam_k"))
            i_ms = k_ms.get("nrc_infer_k", 0.0)
            t_tf = NRC_TRAIN_FLOPS * n_train / (t_ms * 1e-3) / 1e12 if t_ms else None
            i_tf = NRC_FWD_FLOPS * int(st["visible_pixels"]) / (i_ms * 1e-3) / 1e12 if i_ms else None
            out[mode]["mfma"] = {"train_samples_per_frame": n_train, "train_tflops": t_tf, "infer_tflops": i_tf,
                                 "peak_tflops": MFMA_BF16_PEAK_TFLOPS,
                                 "train_frac": t_tf / MFMA_BF16_PEAK_TFLOPS if t_tf else None,
                                 "infer_frac": i_tf / MFMA_BF16_PEAK_TFLOPS if i_tf else None}
    if plain_rays is not None:
        out["neural"]["rays_per_s_equivalent"] = plain_rays / (out["neural"]["frame_ms"] * 1e-3)
        out["neural"]["speedup_vs_plain"] = out["plain"]["frame_ms"] / out["neural"]["frame_ms"]
    for q in (vis_d, rad_d, irr_d):
        ctx.free(q)
    if not args.no_cpu_baseline:
        log("batch: scene export for the CPU baseline")
        scene = ctx.scene_export(params)
        ctx.close()
        log("batch: cpu baseline")
        out["cpu_baseline"] = batch_cpu_baseline(scene, cams, vis_host)
        del scene
    else:
        ctx.close()
    return out


def batch_cpu_baseline(scene, cams, vis_gpu, row_step: int = 8, budget_s: float = 25.0) -> dict:
    """The oracle's culled visibility of every `row_step`-th row of each
    camera of the batched frame (as many cameras as fit in the budget), on
    the host cores. The same rows of the GPU's visibility buffer are compared
    word for word."""
    import oracle  # the CPU baseline leg is the only bench code that uses the oracle

    oscene = oracle.OracleScene(scene)
    threads = min(int(os.environ.get("OMP_NUM_THREADS", "0")) or (os.cpu_count() or 1), 16)
    rows = len(range(0, H, row_step))
    exact, rays = [], 0
    t0 = time.perf_counter()
    for c in range(cams.shape[0]):
        ref, _ = oracle.visibility_culled(oscene, cams[c:c + 1], W, H, row_step=row_step, n_threads=threads)
        exact.append(bool(np.array_equal(ref[::row_step], vis_gpu[c, ::row_step])))
        rays += W * rows
        if time.perf_counter() - t0 > budget_s:
            break
    elapsed = time.perf_counter() - t0
    return {"value": rays / elapsed, "unit": "rays/s", "cores": threads, "kind": "port",
            "sample": f"oracle culled visibility (cull of all {scene.n_clusters} clusters per camera, then every "
                      f"{row_step}th row, {rows} rows x {W}) of {len(exact)} of the {cams.shape[0]} cameras, "
                      f"{elapsed:.1f} s on {threads} threads",
            "visibility_check": {"cameras_checked": len(exact), "row_step": row_step,
                                 "bit_exact": bool(exact) and all(exact), "per_camera": exact}}


def shading_bench(ctx, cam, vis_d, steps: int) -> dict:
    """The shading pass (shading.frag's full-screen pass, lit view) over the
    frame's visibility buffer. Algorithmic bytes per launch: the 8-B word
    read and the 16-B RGBA written per pixel, plus per visible pixel its
    4-B cluster pointer entry, 4-B cluster->object entry and 1-B material."""
    rgba = ctx.alloc(16 * W * H)
    prm = T.shade_params()
    ctx.shade_device(cam, W, H, vis_d, prm, rgba)
    ctx.synchronize()
    ctx.profile(True)
    ctx.profile_reset()
    for _ in range(steps):
        ctx.shade_device(cam, W, H, vis_d, prm, rgba)
    ctx.synchronize()
    launches, total = ctx.profile_stats()["shade_k"]
    ctx.profile(False)
    ctx.free(rgba)
    vis = np.empty((H, W), np.uint64)
    ctx.d2h(vis, vis_d)
    visible = int((vis != T.VISIBILITY_EMPTY).sum())
    alg = 24 * W * H + 9 * visible
    avg_ms = total / launches
    gbs = alg / (avg_ms * 1e-3) / 1e9
    return {"kernel": "shade_k", "kernel_avg_ms": avg_ms, "alg_bytes_per_launch": alg, "achieved_gbs": gbs,
            "frac": gbs / HBM_PEAK_GBS, "view": "lit (tg's default constants)", "visible_pixels": visible}


def svo_visibility_bench(ctx, cam, steps: int) -> dict:
    """SVO primary visibility (the second producer of the visibility word)
    over the 1080p frame: primary rays per second through the SVO."""
    out = ctx.alloc(8 * W * H)
    ctx.svo_visibility_device(cam, W, H, out)
    ctx.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        ctx.svo_visibility_device(cam, W, H, out)
    ctx.synchronize()
    ms = 1e3 * (time.perf_counter() - t0) / steps
    vis = np.empty((H, W), np.uint64)
    ctx.d2h(vis, out)
    ctx.free(out)
    return {"ms": ms, "rays_per_s": W * H / (ms * 1e-3), "hit_pixels": int((vis != T.VISIBILITY_EMPTY).sum())}


def svo_file_bench(ctx, cam, prm, vis_d, rad_d, irr_d) -> dict:
    """The frame's SVO saved to a file and loaded back in place of the build:
    file size, wall times (synchronous calls), and one GI frame after the
    load compared byte for byte with one before it."""
    def frame() -> bytes:
        ctx.frame_device(cam, W, H, prm, vis_d, rad_d, irr_d)
        out = np.empty((H, W, 3), np.float32)
        ctx.d2h(out, irr_d)
        return out.tobytes()

    before = frame()
    with tempfile.TemporaryDirectory(prefix="tgx_svo_", dir="/tmp") as d:
        path = Path(d) / "frame.tgxsvo"
        t0 = time.perf_counter()
        ctx.svo_save(path)
        save_ms = 1e3 * (time.perf_counter() - t0)
        size = path.stat().st_size
        t0 = time.perf_counter()
        ctx.svo_load(path)
        load_ms = 1e3 * (time.perf_counter() - t0)
    info = ctx.svo_info()
    return {"bytes": size, "save_ms": save_ms, "load_ms": load_ms, "nodes": int(info["n_nodes"]),
            "leaves": int(info["n_leaves"]), "device_bytes": ctx.svo_device_bytes()[0],
            "frame_after_load_identical": frame() == before}


def dynamic_bench(ctx, scene, mn, depth: int) -> dict:
    """The frame's scene rebuilt as a dynamic scene, object by object, then
    the first object destroyed (every later pointer shifts) and created
    again: wall time per call, each call synchronous. Then the SVO after
    that edit: tgx_svo_update (reusing the clean leaves of a build before
    the edit) against a full tgx_svo_build of the same scene."""
    objs = scene.objects
    ctx.scene_reserve(objs.size, scene.n_clusters, scene.luts)
    t0 = time.perf_counter()
    for o in range(objs.size):
        ob = objs[o]
        n = int(np.prod(ob["n_cluster_pointers_per_dim"].astype(np.int64)))
        f = int(ob["first_cluster_pointer"])
        cidx = scene.cluster_pointers[f:f + n].astype(np.int64)
        bits = scene.bits.reshape(-1, 16)[cidx]
        mats = scene.materials.reshape(-1, 512)[cidx]
        ctx.object_create(ob["n_cluster_pointers_per_dim"], ob["translation"], ob["rotation"], int(ob["lut_idx"]),
                          bits, mats)
    create_all = time.perf_counter() - t0
    ctx.svo_build(mn, depth)
    ob = objs[0]
    t0 = time.perf_counter()
    ctx.object_destroy(0)
    destroy_s = time.perf_counter() - t0
    n = int(np.prod(ob["n_cluster_pointers_per_dim"].astype(np.int64)))
    cidx = scene.cluster_pointers[:n].astype(np.int64)
    t0 = time.perf_counter()
    ctx.object_create(ob["n_cluster_pointers_per_dim"], ob["translation"], ob["rotation"], int(ob["lut_idx"]),
                      scene.bits.reshape(-1, 16)[cidx], scene.materials.reshape(-1, 512)[cidx])
    create_s = time.perf_counter() - t0
    live = int(ctx.scene_counts()["n_cluster_pointers"])
    t0 = time.perf_counter()
    reused, voxelized = ctx.svo_update()
    update_s = time.perf_counter() - t0
    upd = ctx.svo_export()
    t0 = time.perf_counter()
    ctx.svo_build(mn, depth)
    build_s = time.perf_counter() - t0
    full = ctx.svo_export()
    same = all(np.array_equal(upd[k], full[k]) for k in ("nodes", "bits", "materials", "leaf_n", "list_offsets", "list"))
    if not same:
        raise SystemExit("tgx_svo_update differs from tgx_svo_build after the edit")
    return {"objects": int(objs.size), "cluster_pointers": live, "create_all_ms": 1e3 * create_all,
            "destroy_first_us": 1e6 * destroy_s, "create_us": 1e6 * create_s,
            "svo_update_ms": 1e3 * update_s, "svo_full_build_ms": 1e3 * build_s,
            "svo_leaves_reused": reused, "svo_leaves_voxelized": voxelized, "svo_update_equals_build": same,
            "note": "destroying object 0 shifts every later cluster pointer on the device; the SVO update "
                    "follows that destroy and re-create of object 0"}


def pick_bench(ctx, cam, vis_d, n: int = 4096) -> dict:
    """Pick readback: latency of one synchronous call for 1 and for n pixels."""
    rng = np.random.default_rng(0)
    xy = np.stack([rng.integers(0, W, n), rng.integers(0, H, n)], 1).astype(np.uint32)
    ctx.pick(vis_d, W, H, xy[:1], cam)
    res = {}
    for k in (1, n):
        t0 = time.perf_counter()
        for _ in range(20):
            ctx.pick(vis_d, W, H, xy[:k], cam)
        res[f"us_per_call_{k}"] = 1e6 * (time.perf_counter() - t0) / 20
    return res


COPY_BYTES = 256 << 20  # known-size copy that calibrates the HBM byte counters


def log(msg: str) -> None:
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def probe(args) -> int:
    """Short run for the byte counters: one warm-up and `steps` frames of
    the workload, then a COPY_BYTES device-to-device copy (tgx_copy_k)."""
    params = T.noise_objects_params(seed=args.seed)
    scene = T.generate_scene(params)
    cam = T.generate_camera(params, 0, W, H)
    mn = np.floor(cam[0]["camera"][:3]) - float(1 << (DEPTH - 1))
    prm = T.gi_params(T.seeded_sun(args.seed), spp=SPP, max_distance=1024.0)
    ctx = T.Context(0)
    ctx.upload(scene)
    vis_d, rad_d, irr_d = ctx.alloc(W * H * 8), ctx.alloc(W * H * 12), ctx.alloc(W * H * 12)
    if args.workload == "frame":
        ctx.svo_build(mn, DEPTH)
    for _ in range(1 + args.steps):
        if args.workload == "frame":
            ctx.frame_device(cam, W, H, prm, vis_d, rad_d, irr_d)
        else:
            ctx.visibility_device(cam, W, H, vis_d)
    a, b = ctx.alloc(COPY_BYTES), ctx.alloc(COPY_BYTES)
    ctx.d2d(b, a, COPY_BYTES)
    ctx.synchronize()
    for q in (vis_d, rad_d, irr_d, a, b):
        ctx.free(q)
    ctx.close()
    return 0


def counter_per_dispatch(db: Path, counter: str) -> dict:
    """Mean per-dispatch value of `counter` by kernel name (short name)."""
    con = sqlite3.connect(str(db))
    rows = con.execute("select kernel_name, dispatch_id, sum(value) from counters_collection "
                       "where counter_name = ? group by dispatch_id", (counter,)).fetchall()
    con.close()
    acc: dict = {}
    for name, _, v in rows:
        short = name.split("(")[0].replace("void ", "").split("<")[0].strip()
        acc.setdefault(short, []).append(float(v))
    return {k: sum(v) / len(v) for k, v in acc.items()}


def pmc_pass(counters: list[str], workload: str, steps: int) -> dict | None:
    """One rocprofv3 --pmc pass over a probe run: {counter: {kernel: mean
    per-dispatch value}}, or None if rocprofv3 is unavailable or fails."""
    rp = shutil.which("rocprofv3")
    if rp is None:
        return None
    log(f"rocprofv3 --pmc {' '.join(counters)} pass")
    with tempfile.TemporaryDirectory(prefix="tgx_pmc_", dir="/tmp") as d:
        cmd = [rp, "--pmc", *counters, "-d", d, "-o", "run", "--", sys.executable, str(ROOT / "bench.py"),
               "--probe", "--workload", workload, "--steps", str(steps)]
        proc = subprocess.Popen(cmd, cwd="/tmp", env=dict(os.environ, TMPDIR="/tmp"), stdout=subprocess.DEVNULL,
                                stderr=subprocess.DEVNULL, start_new_session=True)
        try:
            proc.wait(timeout=150)
        except subprocess.TimeoutExpired:
            os.killpg(proc.pid, signal.SIGKILL)
            proc.wait()
            log("pass timed out")
            return None
        dbs = sorted(Path(d).rglob("*.db"))
        if not dbs:
            log(f"pass left no results (exit {proc.returncode})")
            return None
        return {c: counter_per_dispatch(dbs[0], c) for c in counters}


def measure_traffic(workload: str, kernel: str, steps: int = 2) -> dict | None:
    """HBM bytes per launch of `kernel`: rocprofv3 FETCH_SIZE and WRITE_SIZE
    passes (KiB per dispatch) over a probe run, each scaled by what the same
    counter reads for a copy of COPY_BYTES (tgx_copy_k); then the kernel's L2
    hit rate from TCC_HIT / TCC_MISS (one more pass). Runs rocprofv3 as a
    child process; returns None if it is unavailable or fails."""
    got = {}
    for counter in ("FETCH_SIZE", "WRITE_SIZE"):
        r = pmc_pass([counter], workload, steps)
        if r is None:
            return None
        vals = r[counter]
        if kernel not in vals or "tgx_copy_k" not in vals or vals["tgx_copy_k"] <= 0:
            log(f"{counter}: kernels {sorted(vals)} lack {kernel} or the copy")
            return None
        got[counter] = (vals[kernel], vals["tgx_copy_k"])
    f_k, f_c = got["FETCH_SIZE"]
    w_k, w_c = got["WRITE_SIZE"]
    fetch_scale = COPY_BYTES / 1024.0 / f_c
    write_scale = COPY_BYTES / 1024.0 / w_c
    out = {"bytes": 1024.0 * (f_k * fetch_scale + w_k * write_scale),
           "fetch_kib": f_k, "write_kib": w_k, "copy_fetch_kib": f_c, "copy_write_kib": w_c,
           "copy_bytes": COPY_BYTES, "fetch_scale": fetch_scale, "write_scale": write_scale}
    r = pmc_pass(["TCC_HIT_sum", "TCC_MISS_sum"], workload, steps)
    if r is not None and kernel in r["TCC_HIT_sum"] and kernel in r["TCC_MISS_sum"]:
        hit, miss = r["TCC_HIT_sum"][kernel], r["TCC_MISS_sum"][kernel]
        out["l2_hits"], out["l2_misses"] = hit, miss
        out["l2_hit_rate"] = hit / (hit + miss) if hit + miss > 0 else None
    # lanes active per VALU instruction: thread-cycles over 64 x instruction-cycles
    r = pmc_pass(["SQ_THREAD_CYCLES_VALU", "SQ_ACTIVE_INST_VALU", "SQ_INSTS_VALU"], workload, steps)
    if r is not None and all(kernel in r[c] for c in r):
        tc, ai = r["SQ_THREAD_CYCLES_VALU"][kernel], r["SQ_ACTIVE_INST_VALU"][kernel]
        out["valu"] = {"thread_cycles": tc, "active_inst_cycles": ai, "insts": r["SQ_INSTS_VALU"][kernel],
                       "lane_activity": tc / (64.0 * ai) if ai > 0 else None}
    return out


def copy_bandwidth(ctx, nbytes: int = COPY_BYTES, reps: int = 10) -> float:
    """Attainable HBM bandwidth on this box: a device-to-device copy by the
    library's copy kernel, read + written bytes over the wall time."""
    a, b = ctx.alloc(nbytes), ctx.alloc(nbytes)
    try:
        ctx.d2d(b, a, nbytes)
        ctx.synchronize()
        t0 = time.perf_counter()
        for _ in range(reps):
            ctx.d2d(b, a, nbytes)
        ctx.synchronize()
        return 2.0 * nbytes * reps / (time.perf_counter() - t0) / 1e9
    finally:
        ctx.free(a)
        ctx.free(b)


def cpu_baseline(scene, params, cam, mn, prm, workload: str, budget_s: float = 12.0) -> dict:
    import oracle  # the CPU baseline leg is the only bench code that uses the oracle

    oscene = oracle.OracleScene(scene)
    threads = int(os.environ.get("OMP_NUM_THREADS", "0")) or (os.cpu_count() or 1)
    threads = min(threads, 16)
    if workload == "visibility":
        t0 = time.perf_counter()
        rays, ci = 0, 0
        while time.perf_counter() - t0 < budget_s:
            c = T.generate_camera(params, ci, W, H)
            oracle.visibility_culled(oscene, c, W, H, n_threads=threads)
            rays += W * H
            ci += 1
        rate = rays / (time.perf_counter() - t0)
        return {"value": rate, "unit": "rays/s", "cores": threads, "kind": "port",
                "sample": f"oracle culled visibility, {ci} full 1920x1080 frames on {threads} threads"}
    # Frame workload: whole frames (visibility, then GI of every pixel with
    # the frame's own bounce samples), repeated until the budget is spent.
    t0 = time.perf_counter()
    svo = oracle.OracleSvo(oscene, mn, DEPTH)
    t_build = time.perf_counter() - t0
    frames, rays, t_vis, t_gi = 0, 0, 0.0, 0.0
    vis = None
    while t_vis + t_gi < budget_s:
        t0 = time.perf_counter()
        vis, _ = oracle.visibility_culled(oscene, cam, W, H, n_threads=threads)
        t_vis += time.perf_counter() - t0
        p = prm.copy()
        p["frame"] = frames
        t0 = time.perf_counter()
        _, _, cnt = svo.gi(cam, W, H, vis, p)
        t_gi += time.perf_counter() - t0
        rays += W * H + int(cnt[1] + cnt[3])
        frames += 1
    # one thread, every 16th row of one frame (visibility and GI)
    t0 = time.perf_counter()
    vis1, _ = oracle.visibility_culled(oscene, cam, W, H, row_step=16, n_threads=1)
    _, _, cnt1 = svo.gi(cam, W, H, vis1, prm, row_step=16)
    t1 = time.perf_counter() - t0
    rays1 = W * ((H + 15) // 16) + int(cnt1[1] + cnt1[3])
    cpu_model = ""
    try:
        cpu_model = next(l.split(":", 1)[1].strip() for l in open("/proc/cpuinfo") if l.startswith("model name"))
    except (OSError, StopIteration):
        pass
    # Algorithmic bytes of the timed GPU frame (same camera, visibility and
    # GI parameters): unique 64-B lines the oracle's execution of the frame
    # reads, per part, plus the bytes it writes.
    t0 = time.perf_counter()
    alg = svo.gi_alg_bytes(cam, W, H, vis, prm)
    alg["oracle_s"] = time.perf_counter() - t0
    return {"value": rays / (t_vis + t_gi), "unit": "rays/s", "cores": threads, "kind": "port",
            "alg_bytes": alg,
            "sample": f"{frames} whole 1920x1080 frames of the oracle (visibility {t_vis:.2f} s, GI at {SPP} spp "
                      f"{t_gi:.2f} s, {rays} primary+bounce+shadow rays) on {threads} threads; oracle SVO build "
                      f"{t_build:.2f} s",
            "single_thread": {"value": rays1 / t1, "unit": "rays/s",
                              "sample": f"every 16th row of one frame, visibility and GI, {rays1} rays, {t1:.2f} s"},
            "cpu_model": cpu_model, "host_cpus": os.cpu_count()}


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--seed", type=int, default=1)
    ap.add_argument("--workload", choices=("frame", "visibility", "batch"), default="frame")
    ap.add_argument("--batch-pretrain", type=int, default=60, help="untimed (training) replays of the batched cache frame")
    ap.add_argument("--batch-steps", type=int, default=5, help="timed replays of each batched frame graph")
    ap.add_argument("--no-batch", action="store_true", help="skip the batched 16-camera frame")
    ap.add_argument("--batch-no-graph", action="store_true",
                    help="time uncaptured batched frames (the same launches) instead of graph replays")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-neural", action="store_true", help="skip the neural-cache measurement")
    ap.add_argument("--neural-pretrain", type=int, default=12800,
                    help="untimed cache training frames of the static view before the error is read")
    ap.add_argument("--no-traffic", action="store_true", help="skip the rocprofv3 byte-counter passes")
    ap.add_argument("--probe", action="store_true", help=argparse.SUPPRESS)
    args = ap.parse_args()
    if args.probe:
        return probe(args)
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU: this build has no multi-GPU path (each GPU would render whole "
              "cameras or screen tiles of its own; see the sharding design). Run with --gpus 1.", file=sys.stderr)
        return 2

    if args.workload == "batch":
        b = batch_bench(args, args.steps, args.batch_pretrain)
        line = {
            "metric": "rays/s (primary DDA + SVO bounce, 1080p); neural vs plain GI at <=2% rel-MSE",
            "value": b["plain"]["rays_traced_per_s"], "unit": "rays/s", "n_gpus": 1, "steps": args.steps,
            "warmup": args.batch_pretrain, "ms_per_step": b["plain"]["frame_ms"], "higher_is_better": True,
            "scaling": "weak", "vs_baseline": None, "dtype": "f32", "data": "synthetic",
            "config": {"workload": BATCH_WORKLOAD, "cameras": BATCH_CAMS, "width": W, "height": H},
            "roofline": b["plain"]["roofline"], "batch": b,
        }
        if "cpu_baseline" in b:
            line["cpu_baseline"] = b["cpu_baseline"]
        print(json.dumps(line), flush=True)
        return 0

    params = T.noise_objects_params(seed=args.seed)
    scene = T.generate_scene(params)
    cam = T.generate_camera(params, 0, W, H)
    mn = np.floor(cam[0]["camera"][:3]) - float(1 << (DEPTH - 1))
    prm = T.gi_params(T.seeded_sun(args.seed), spp=SPP, max_distance=1024.0)
    ctx = T.Context(0)
    ctx.upload(scene)
    t0 = time.perf_counter()
    ctx.svo_build(mn, DEPTH)
    svo_build_ms = 1e3 * (time.perf_counter() - t0)
    info = ctx.svo_info()
    vis_d = ctx.alloc(W * H * 8)
    rad_d = ctx.alloc(W * H * 12)
    irr_d = ctx.alloc(W * H * 12)

    def step():
        if args.workload == "frame":
            ctx.frame_device(cam, W, H, prm, vis_d, rad_d, irr_d)
        else:
            ctx.visibility_device(cam, W, H, vis_d)

    for _ in range(args.warmup):
        step()
    ctx.synchronize()

    ctx.profile(True)
    ctx.profile_reset()
    ctx.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        step()
    ctx.synchronize()
    elapsed = time.perf_counter() - t0
    prof = ctx.profile_stats()
    ctx.profile(False)
    vstats = ctx.last_stats()

    ms_per_step = 1e3 * elapsed / args.steps
    primary = W * H
    kernels_ms = {k: v[1] / args.steps for k, v in prof.items()}
    if args.workload == "frame":
        gst = ctx.gi_stats()
        rays_per_frame = primary + int(gst["bounce_rays"]) + int(gst["shadow_rays"])
        dominant = max(("gi_trace_k", "vis_traverse"), key=lambda k: kernels_ms.get(k, 0.0))
        # per-launch bytes: one trace launch per frame (bounce + shadow rays)
        alg = (trace_alg_bytes(gst, info) * args.steps // prof["gi_trace_k"][0] if dominant == "gi_trace_k"
               else vis_alg_bytes(vstats, scene.objects.size, primary))
    else:
        gst = None
        rays_per_frame = primary
        dominant = "vis_traverse"
        alg = vis_alg_bytes(vstats, scene.objects.size, primary)
    launches, total_ms = prof[dominant]
    avg_ms = total_ms / launches
    achieved = alg / (avg_ms * 1e-3) / 1e9
    vis_ms = sum(v for k, v in kernels_ms.items() if k.startswith("vis_"))

    line = {
        "metric": "rays/s (primary DDA + SVO bounce, 1080p); neural vs plain GI at <=2% rel-MSE",
        "value": rays_per_frame * args.steps / elapsed,
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
            "workload": ("primary visibility + one-bounce SVO GI" if args.workload == "frame" else "primary visibility")
                        + ", 512 noise-filled objects of 1-8 clusters per axis in a 512^3 world, 1920x1080, camera at "
                          "the world centre (seeded yaw/pitch), bit-exact 64-bit visibility"
                        + (f", depth-{DEPTH} SVO around the camera, {SPP} spp cosine bounce rays, range 1024 voxels, "
                           "seeded sun, one shadow ray per sun-facing bounce hit" if args.workload == "frame" else ""),
            "objects": int(scene.objects.size), "clusters": scene.n_clusters, "width": W, "height": H,
            "cameras": 1, "seed": args.seed, "spp": SPP if args.workload == "frame" else 0,
            "rays_per_frame": rays_per_frame,
            "rays_counted": "primary (one per pixel) + bounce + shadow" if args.workload == "frame" else "primary",
        },
        "roofline": {
            "bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS, "traffic": None,
            "kernel": dominant, "kernel_avg_ms": avg_ms, "alg_bytes_per_launch": alg,
        },
        "kernels_ms_per_step": kernels_ms,
        "primary_visibility": {"ms": vis_ms, "rays_per_s": primary / (vis_ms * 1e-3) if vis_ms else None,
                               "clusters_binned": int(vstats.clusters_binned),
                               "tile_entries": int(vstats.tile_entries)},
        "svo": svo_memory(ctx, info, {"build_ms": svo_build_ms, "nodes": int(info["n_nodes"]),
                                       "leaves": int(info["n_leaves"])}),
    }
    log(f"timed {args.steps} steps: {ms_per_step:.3f} ms/step")
    line["shading"] = shading_bench(ctx, cam, vis_d, args.steps)
    line["svo_visibility"] = svo_visibility_bench(ctx, cam, args.steps)
    line["pick"] = pick_bench(ctx, cam, vis_d)
    if args.workload == "frame":
        log("svo file save / load")
        line["svo_file"] = svo_file_bench(ctx, cam, prm, vis_d, rad_d, irr_d)
    if gst is not None:
        line["gi"] = {k: int(gst[k]) for k in gst.dtype.names}
        if not args.no_neural:
            log("neural cache measurement")
            line["neural"] = neural_bench(ctx, cam, T.seeded_sun(args.seed), vis_d, rad_d, irr_d, args.steps,
                                          args.neural_pretrain, ms_per_step, rays_per_frame)
    for p in (vis_d, rad_d, irr_d):
        ctx.free(p)
    # the roofline's attainable counterpart: a streaming copy on this box
    line["roofline"]["attainable_copy_gbs"] = copy_bandwidth(ctx)
    if args.workload == "frame":
        log("dynamic scene edits")
        line["dynamic"] = dynamic_bench(ctx, scene, mn, DEPTH)
    ctx.close()
    if not args.no_batch and args.workload == "frame":
        log("batched 16-camera frame")
        line["batch"] = batch_bench(args, args.batch_steps, args.batch_pretrain)
        line["batch"]["workload"] = BATCH_WORKLOAD
    if not args.no_traffic:
        tr = measure_traffic(args.workload, dominant)
        if tr is not None:
            line["roofline"]["traffic"] = tr["bytes"]
            line["roofline"]["traffic_detail"] = tr
            line["roofline"]["l2_hit_rate"] = tr.get("l2_hit_rate")
    if not args.no_cpu_baseline:
        log("cpu baseline")
        cb = cpu_baseline(scene, params, cam, mn, prm, args.workload)
        alg = cb.pop("alg_bytes", None)
        line["cpu_baseline"] = cb
        if alg is not None and line["roofline"]["kernel"] == "gi_trace_k":
            # The trace launch's bytes by the oracle's line count: the SVO
            # lines (nodes, leaf data, voxel bit blocks, leaf materials) and
            # LUT lines its bounce and shadow rays read. The frame's surface
            # reads and radiance writes belong to the surface and resolve
            # kernels; the whole frame's figure is reported beside it.
            r = line["roofline"]
            r["alg_bytes_closed_form"] = r["alg_bytes_per_launch"]
            r["alg_bytes_per_launch"] = alg["trace"]
            r["achieved"] = alg["trace"] / (r["kernel_avg_ms"] * 1e-3) / 1e9
            r["frac"] = r["achieved"] / HBM_PEAK_GBS
            r["alg_bytes_source"] = ("oracle: 64 B x unique lines read by the trace of the frame's bounce and "
                                     "shadow rays (tgxo_gi_alg_bytes)")
            r["frame_alg_bytes"] = alg
            r["frame_frac"] = alg["total"] / (line["ms_per_step"] * 1e-3) / 1e9 / HBM_PEAK_GBS
    print(json.dumps(line), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
