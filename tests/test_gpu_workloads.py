"""The two largest workloads at their full sizes on the GPU.

- Batched frame: 4096 noise-filled objects of 1-32 clusters per axis in a
  4096^3 world with 64 LUTs, generated on the device; a depth-12 SVO; 16
  seeded cameras at 1920x1080 rendered as one captured frame graph, plain
  GI at 4 spp. Every camera's visibility is compared word for word with the
  oracle's culled visibility on every 8th row. The neural-cache form of the
  same frame (fallback against the tracer's variance) then replays and its
  fallback traces only part of the visible pixels once trained.
- Neural cache at 1920x1080 on the 512-object scene, with fallback: the
  first frame traces every visible pixel (nothing learnt), later frames
  trace fewer, and after 200 training frames the frame's irradiance is far
  closer to a 256-spp reference than a 4-spp plain frame.

Each test uses a context of its own, closed at its end, so the large
buffers do not stay with the session context."""
from __future__ import annotations

import os
import sys

import numpy as np
import pytest

import oracle as O
import tg_amd as T
from oracle import nrc

pytestmark = pytest.mark.gpu

W, H = 1920, 1080


def say(msg: str) -> None:
    print(f"[gpu] {msg}", file=sys.stderr, flush=True)


def test_batched_frame_16_cameras_depth12_visibility_every_8th_row():
    params = T.noise_objects_params(seed=1, n_objects=4096, max_extent=32, world=4096.0, n_luts=64)
    n_cams, depth, row_step = 16, 12, 8
    cams = np.concatenate([T.generate_camera(params, i, W, H) for i in range(n_cams)])
    mn = np.floor(cams[0]["camera"][:3]) - float(1 << (depth - 1))
    prm = T.gi_params(T.seeded_sun(1), spp=4, max_distance=1024.0)
    ctx = T.Context(int(os.environ.get("TGX_DEVICE", "0")))
    n_px = n_cams * W * H
    bufs = []
    try:
        say("batched workload: device scene generation (4096 objects) and depth-12 SVO build")
        ctx.generate_scene_device(params)
        ctx.svo_build(mn, depth)
        info = ctx.svo_info()
        assert int(info["n_leaves"]) > 100_000
        bufs = [ctx.alloc(8 * n_px), ctx.alloc(12 * n_px), ctx.alloc(12 * n_px)]
        say("batched workload: captured plain frame, 16 cameras at 1920x1080")
        g = ctx.frame_graph(cams, W, H, prm, *bufs)
        try:
            g.launch()
            st = g.stats()
        finally:
            g.destroy()
        vis = np.empty((n_cams, H, W), np.uint64)
        rad = np.empty((n_cams, H, W, 3), np.float32)
        ctx.d2h(vis, bufs[0])
        ctx.d2h(rad, bufs[1])
        assert int(st["primary_rays"]) == n_px
        assert int(st["visible_pixels"]) > 0.3 * n_px
        assert int(st["bounce_rays"]) == 4 * int(st["visible_pixels"]) and int(st["shadow_rays"]) > 0
        assert np.isfinite(rad).all() and rad.min() >= 0.0 and rad.sum() > 0
        say("batched workload: neural-cache frame graph with fallback, 30 replays")
        ctx.nrc_create(T.nrc_config(origin=(0.0, 0.0, 0.0), scale=1.0 / 4096, train_samples=1 << 16,
                                    train_batch=1 << 12, fallback_threshold=16.0, fallback_min_samples=0,
                                    flags=T.NRC_FLAG_FALLBACK_VS_TRACER, ema_decay=0.999))
        g = ctx.frame_graph(cams, W, H, prm, *bufs, flags=T.FRAME_NEURAL)
        try:
            for _ in range(30):
                g.launch()
            nst = g.stats()
        finally:
            g.destroy()
        ctx.d2h(rad, bufs[1])
        assert int(nst["visible_pixels"]) == int(st["visible_pixels"])
        assert 0 < int(nst["fallback_pixels"]) < int(nst["visible_pixels"])
        assert int(nst["train_rays"]) >= n_cams * (1 << 16)
        assert np.isfinite(rad).all() and rad.sum() > 0
        say("batched workload: scene export for the oracle")
        scene = ctx.scene_export(params)
    finally:
        for q in bufs:
            ctx.free(q)
        ctx.close()
    oscene = O.OracleScene(scene)
    threads = min(int(os.environ.get("OMP_NUM_THREADS", "0")) or (os.cpu_count() or 1), 16)
    # Visible pixels are words with depth < 1 whose hit voxel the surface
    # pass re-intersects; words at depth exactly 1 are hits of the buffer but
    # not surfaces.
    n_surf = sum(O.gi_surface_count(oscene, cams[c:c + 1], W, H, vis[c], threads) for c in range(n_cams))
    say(f"batched workload: {int(st['visible_pixels'])} visible pixels, oracle surfaces {n_surf}, "
        f"non-empty words {int((vis != T.VISIBILITY_EMPTY).sum())}")
    assert int(st["visible_pixels"]) == n_surf
    for c in range(n_cams):
        say(f"batched workload: oracle visibility, camera {c}, every {row_step}th row")
        ref, _ = O.visibility_culled(oscene, cams[c:c + 1], W, H, row_step=row_step, n_threads=threads)
        bad = int((ref[::row_step] != vis[c, ::row_step]).sum())
        assert bad == 0, f"camera {c}: {bad} visibility words differ from the oracle"


def reference_irradiance(ctx, cam, w, h, sun, frames: int = 64, first: int = 1000) -> np.ndarray:
    """256-spp irradiance: 64 independent 4-spp frames accumulated in float64."""
    acc = np.zeros((h, w, 3), np.float64)
    for f in range(frames):
        _, _, irr = ctx.frame(cam, w, h, T.gi_params(sun, spp=4, frame=first + f))
        acc += irr
    return acc / frames


def test_neural_cache_1080p_with_fallback():
    p = T.noise_objects_params(seed=1)
    sc = T.generate_scene(p)
    cam = T.generate_camera(p, 0, W, H)
    sun = T.seeded_sun(1)
    ctx = T.Context(int(os.environ.get("TGX_DEVICE", "0")))
    try:
        ctx.upload(sc)
        ctx.svo_build(np.floor(cam[0]["camera"][:3]) - 1024.0, 11)
        say("cache 1080p: 256-spp reference")
        ref = reference_irradiance(ctx, cam, W, H, sun)
        vis, _, plain = ctx.frame(cam, W, H, T.gi_params(sun, spp=4, frame=7))
        mask = vis != T.VISIBILITY_EMPTY
        e_plain = nrc.rel_mse(plain[mask], ref[mask])
        ctx.nrc_create(T.nrc_config(origin=(0, 0, 0), scale=1 / 512, train_samples=1 << 16, train_batch=1 << 12,
                                    fallback_threshold=1.0, fallback_min_samples=1,
                                    flags=T.NRC_FLAG_FALLBACK_VS_TRACER))
        prm = T.gi_params(sun, spp=4, frame=0)
        ctx.nrc_frame(cam, W, H, prm)
        first = ctx.nrc_stats()
        # visible pixels: words with a surface (a word at depth exactly 1 is none)
        n_surf = O.gi_surface_count(O.OracleScene(sc), cam, W, H, vis)
        assert first["fallback_pixels"] == first["visible_pixels"] == n_surf
        assert 0 <= int(mask.sum()) - n_surf < 100
        say("cache 1080p: 200 training frames with fallback")
        for _ in range(200):
            _, _, irr = ctx.nrc_frame(cam, W, H, prm)
        st = ctx.nrc_stats()
        e_nrc = nrc.rel_mse(irr[mask], ref[mask])
        say(f"cache 1080p: rel-MSE vs 256 spp: plain 4 spp {e_plain:.4f}, cache frame {e_nrc:.4f}, "
            f"fallback {int(st['fallback_pixels'])} of {int(st['visible_pixels'])} pixels")
        assert 0 < st["fallback_pixels"] < st["visible_pixels"] and st["fallback_rays"] > 0
        assert st["adam_steps"] == 201 * 16
        assert e_nrc < 0.1 * e_plain
    finally:
        ctx.close()
