"""Batched multi-camera frames: the device scene generator against the host
generator, captured frame graphs against uncaptured frames and the
visibility oracle, the context's buffer lock while a graph exists, and the
autotuner (whose candidates must not change any result).

The scene has the batched workload's shape at a size the oracle finishes
in seconds: objects of 1..32 clusters per axis, value-noise solids, 64
LUTs, several seeded cameras at the world centre."""
from __future__ import annotations

import os
import sys

import numpy as np
import pytest

import oracle as O
import tg_amd as T

pytestmark = pytest.mark.gpu

W, H = 320, 180
N_CAMS = 4
DEPTH = 10


def say(msg: str) -> None:
    print(f"[gpu] {msg}", file=sys.stderr, flush=True)


def batch_params(n_objects: int = 12):
    return T.noise_objects_params(seed=7, n_objects=n_objects, max_extent=32, world=512.0, n_luts=64)


@pytest.fixture(scope="module")
def batch_scene():
    p = batch_params()
    sc = T.generate_scene(p)
    cams = np.concatenate([T.generate_camera(p, i, W, H) for i in range(N_CAMS)])
    mn = np.floor(cams[0]["camera"][:3]) - float(1 << (DEPTH - 1))
    return p, sc, cams, mn


class Buffers:
    def __init__(self, ctx, n_cams: int, w: int, h: int):
        self.ctx, self.n, self.w, self.h = ctx, n_cams, w, h
        self.vis = ctx.alloc(8 * n_cams * w * h)
        self.rad = ctx.alloc(12 * n_cams * w * h)
        self.irr = ctx.alloc(12 * n_cams * w * h)

    def read(self):
        vis = np.empty((self.n, self.h, self.w), np.uint64)
        rad = np.empty((self.n, self.h, self.w, 3), np.float32)
        irr = np.empty((self.n, self.h, self.w, 3), np.float32)
        self.ctx.d2h(vis, self.vis)
        self.ctx.d2h(rad, self.rad)
        self.ctx.d2h(irr, self.irr)
        return vis, rad, irr

    def free(self):
        for q in (self.vis, self.rad, self.irr):
            self.ctx.free(q)


def setup_scene(ctx, p, mn):
    say("device scene generation + SVO build")
    ctx.generate_scene_device(p)
    ctx.svo_build(mn, DEPTH)


@pytest.mark.parametrize("kind", ["noise", "single"])
def test_device_scene_equals_host_scene(ctx, kind):
    p = batch_params(20) if kind == "noise" else T.single_object_params(3)
    say(f"device generator vs host generator ({kind})")
    ctx.generate_scene_device(p)
    got = ctx.scene_export(p)
    ref = T.generate_scene(p)
    for name in ("objects", "cluster_pointers", "cluster_idx_to_object", "bits", "materials", "luts"):
        a, b = getattr(got, name), getattr(ref, name)
        assert a.shape == b.shape and a.tobytes() == b.tobytes(), name
    if kind == "noise":
        assert 0.3 < ref.solid_fraction() < 0.5
        n = ref.objects["n_cluster_pointers_per_dim"]
        assert n.min() >= 1 and n.max() <= 32 and n.max() > 8


def test_frame_graph_plain_matches_uncaptured_and_oracle(ctx, batch_scene):
    p, sc, cams, mn = batch_scene
    setup_scene(ctx, p, mn)
    prm = T.gi_params(T.seeded_sun(4), spp=2, frame=3)
    buf = Buffers(ctx, N_CAMS, W, H)
    try:
        say(f"frame graph: {N_CAMS} cameras {W}x{H}, plain GI")
        g = ctx.frame_graph(cams, W, H, prm, buf.vis, buf.rad, buf.irr)
        try:
            g.launch()
            st1 = g.stats()
            vis1, rad1, irr1 = buf.read()
            g.launch()
            st2 = g.stats()
            vis2, rad2, irr2 = buf.read()
        finally:
            g.destroy()
        # plain GI draws the samples of prm.frame: replays are identical
        assert vis1.tobytes() == vis2.tobytes() and rad1.tobytes() == rad2.tobytes()
        assert irr1.tobytes() == irr2.tobytes() and tuple(st1) == tuple(st2)
        osc = O.OracleScene(sc)
        visible = bounce = shadow = hits = 0
        for c in range(N_CAMS):
            say(f"uncaptured frame + oracle visibility, camera {c}")
            v, r, i = ctx.frame(cams[c:c + 1], W, H, prm)
            gs = ctx.gi_stats()
            ref_vis, _ = O.visibility_culled(osc, cams[c:c + 1], W, H)
            assert np.array_equal(vis1[c], ref_vis), f"camera {c}: visibility differs from the oracle"
            assert np.array_equal(vis1[c], v)
            assert rad1[c].tobytes() == r.tobytes(), f"camera {c}: radiance differs from the uncaptured frame"
            assert irr1[c].tobytes() == i.tobytes()
            visible += int(gs["visible_pixels"])
            bounce += int(gs["bounce_rays"])
            shadow += int(gs["shadow_rays"])
            hits += int(gs["bounce_hits"])
        assert st1["primary_rays"] == N_CAMS * W * H
        assert st1["visible_pixels"] == visible == int((vis1 != T.VISIBILITY_EMPTY).sum())
        assert (st1["bounce_rays"], st1["shadow_rays"], st1["bounce_hits"]) == (bounce, shadow, hits)
        assert st1["train_rays"] == 0 and visible > 0.05 * N_CAMS * W * H and rad1.sum() > 0
    finally:
        buf.free()


def test_graph_locks_context_buffers(batch_scene):
    """While a graph exists, calls that would grow a context buffer fail. A
    context of its own, so no earlier test has grown the buffers already."""
    p, sc, cams, mn = batch_scene
    ctx = T.Context(int(os.environ.get("TGX_DEVICE", "0")))
    setup_scene(ctx, p, mn)
    prm = T.gi_params(T.seeded_sun(4), spp=1)
    buf = Buffers(ctx, 2, W, H)
    try:
        g = ctx.frame_graph(cams[:2], W, H, prm, buf.vis, buf.rad, buf.irr)
        try:
            with pytest.raises(T.TgxError):
                ctx.svo_build(mn, DEPTH)
            with pytest.raises(T.TgxError):
                ctx.nrc_create(T.nrc_config())
            big = Buffers(ctx, 1, 4 * W, 4 * H)
            try:
                with pytest.raises(T.TgxError):
                    ctx.frame_device(cams[0:1], 4 * W, 4 * H, prm, big.vis, big.rad, big.irr)
            finally:
                big.free()
            g.launch()
            assert g.stats()["primary_rays"] == 2 * W * H
        finally:
            g.destroy()
        ctx.svo_build(mn, DEPTH)  # unlocked again
    finally:
        buf.free()
        ctx.close()


@pytest.mark.parametrize("nrc_flags", [0, T.NRC_FLAG_FALLBACK_VS_TRACER, T.NRC_FLAG_NO_FALLBACK,
                                       T.NRC_FLAG_ALL_FALLBACK])
def test_frame_graph_neural_matches_uncaptured(ctx, batch_scene, nrc_flags):
    """A captured multi-camera cache frame (training targets of every camera
    in one launch, fallback pixels of every camera in another) equals the
    cache frames run camera by camera, outputs and weights bit for bit, for
    each fallback mode."""
    p, sc, cams, mn = batch_scene
    setup_scene(ctx, p, mn)
    prm = T.gi_params(T.seeded_sun(4), spp=1, frame=5)
    n_cams, replays = 3, 2
    cfg = T.nrc_config(origin=tuple(mn), scale=1.0 / (1 << DEPTH), train_samples=4096, train_batch=2048, seed=9,
                       flags=nrc_flags)
    steps_per_frame = n_cams * 4096 // 2048
    buf = Buffers(ctx, n_cams, W, H)
    try:
        ctx.nrc_create(cfg)
        say(f"frame graph: {n_cams} cameras {W}x{H}, neural cache, {replays} replays")
        g = ctx.frame_graph(cams[:n_cams], W, H, prm, buf.vis, buf.rad, buf.irr, flags=T.FRAME_NEURAL)
        try:
            for _ in range(replays):
                g.launch()
            st = g.stats()
            vis_g, rad_g, irr_g = buf.read()
            assert ctx.nrc_stats()["adam_steps"] == (1 + replays) * steps_per_frame
            w_g = ctx.nrc_params()
        finally:
            g.destroy()
        assert st["train_rays"] > 0 and st["visible_pixels"] == int((vis_g != T.VISIBILITY_EMPTY).sum())
        # the same frames uncaptured, camera by camera, from a fresh cache
        ctx.nrc_create(cfg)
        one = Buffers(ctx, 1, W, H)
        try:
            for f in range(1 + replays):
                for c in range(n_cams):
                    ctx.nrc_frame_device(cams[c:c + 1], W, H, prm, one.vis, one.rad, one.irr)
                    if f == replays:
                        v, r, i = one.read()
                        assert np.array_equal(v[0], vis_g[c])
                        assert r[0].tobytes() == rad_g[c].tobytes(), f"camera {c}: cache radiance differs"
                        assert i[0].tobytes() == irr_g[c].tobytes()
        finally:
            one.free()
        assert ctx.nrc_params().tobytes() == w_g.tobytes()
        assert ctx.nrc_stats()["adam_steps"] == (1 + replays) * steps_per_frame
    finally:
        buf.free()


def test_autotune_keeps_results(ctx, batch_scene):
    p, sc, cams, mn = batch_scene
    setup_scene(ctx, p, mn)
    prm = T.gi_params(T.seeded_sun(4), spp=2, frame=1)
    ctx.set_tuning(trace_sort=0, trace_refill=16, trace_blocks_per_cu=2, trace_vote=0)
    buf = Buffers(ctx, 2, W, H)
    try:
        g = ctx.frame_graph(cams[:2], W, H, prm, buf.vis, buf.rad, buf.irr)
        g.launch()
        before = buf.read()
        g.destroy()
        say("autotune (coordinate descent over trace form, skip, workgroups, DDA, refill, queue chunk, cull "
            "blocks, tile edge), 2 cameras")
        best = ctx.autotune(cams[:2], W, H, prm, buf.vis, buf.rad, buf.irr, frames=1)
        assert tuple(ctx.tuning()[0]) == tuple(best[0])
        assert best["trace_blocks_per_cu"][0] in (1, 2) and best["trace_refill"][0] in (4, 5, 6, 8, 12, 16)
        assert best["trace_chunk"][0] in (0, 32, 64)
        assert best["trace_regroup"][0] in (0, 1, 2, 4) and best["trace_skip"][0] in (0, 1)
        assert best["vis_setup_blocks"][0] in (256, 1024, 2048, 4096) and best["vis_tile"][0] in (8, 16, 32)
        assert best["trace_dda"][0] in (0, 1) and best["trace_node"][0] in (1, 2, 3, 4)
        assert best["trace_cross"][0] in (0, 24, 32, 40, 48)
        assert best["trace_sort"][0] == 0 and best["trace_vote"][0] == 0
        # every form the tuner tries gives the same frame
        for form in ((1, 0, 8), (2, 1, 32), (0, 0, 16)):
            ctx.set_tuning(trace_regroup=form[0], trace_skip=form[1], vis_setup_blocks=256, vis_tile=form[2])
            g = ctx.frame_graph(cams[:2], W, H, prm, buf.vis, buf.rad, buf.irr)
            g.launch()
            got = buf.read()
            g.destroy()
            for a, b in zip(before, got):
                assert a.tobytes() == b.tobytes(), f"form {form}"
        ctx.set_tuning(**{k: int(best[k][0]) for k in best.dtype.names})
        g = ctx.frame_graph(cams[:2], W, H, prm, buf.vis, buf.rad, buf.irr)
        g.launch()
        after = buf.read()
        g.destroy()
        for a, b in zip(before, after):
            assert a.tobytes() == b.tobytes()
    finally:
        ctx.set_tuning(trace_sort=0, trace_refill=5, trace_blocks_per_cu=2, trace_vote=0, trace_regroup=0,
                       trace_skip=0, vis_setup_blocks=2048, trace_dda=1, vis_tile=16, trace_cross=40,
                       trace_node=3, trace_chunk=32)
        buf.free()


def test_frame_graph_joint_training(ctx, batch_scene):
    """Joint training (TGX_NRC_FLAG_JOINT_TRAINING): a captured 3-camera
    cache frame equals the same frames run uncaptured with tgx_frame_batch,
    outputs and weights bit for bit; a frame takes train_samples /
    train_batch Adam steps, each on train_batch samples of every camera."""
    p, sc, cams, mn = batch_scene
    setup_scene(ctx, p, mn)
    prm = T.gi_params(T.seeded_sun(4), spp=1, frame=5)
    n_cams, replays = 3, 2
    cfg = T.nrc_config(origin=tuple(mn), scale=1.0 / (1 << DEPTH), train_samples=4096, train_batch=2048, seed=9,
                       flags=T.NRC_FLAG_FALLBACK_VS_TRACER | T.NRC_FLAG_JOINT_TRAINING)
    steps_per_frame = 4096 // 2048
    buf = Buffers(ctx, n_cams, W, H)
    try:
        ctx.nrc_create(cfg)
        say(f"joint training: {n_cams} cameras {W}x{H}, captured, {replays} replays")
        g = ctx.frame_graph(cams[:n_cams], W, H, prm, buf.vis, buf.rad, buf.irr, flags=T.FRAME_NEURAL)
        try:
            for _ in range(replays):
                g.launch()
            st = g.stats()
            out_g = buf.read()
            w_g = ctx.nrc_params()
            assert ctx.nrc_stats()["adam_steps"] == (1 + replays) * steps_per_frame
        finally:
            g.destroy()
        assert st["train_rays"] > 0
        say("joint training: the same frames uncaptured")
        ctx.nrc_create(cfg)
        for _ in range(1 + replays):
            ctx.frame_batch(cams[:n_cams], W, H, prm, buf.vis, buf.rad, buf.irr, flags=T.FRAME_NEURAL)
        out_b = buf.read()
        for a, b in zip(out_g, out_b):
            assert a.tobytes() == b.tobytes()
        assert ctx.nrc_params().tobytes() == w_g.tobytes()
        assert ctx.nrc_stats()["adam_steps"] == (1 + replays) * steps_per_frame
    finally:
        buf.free()


def test_joint_training_learns_every_camera(ctx):
    """Two cameras trained jointly (2^16 targets per camera and frame, 16
    Adam steps of 2 x 4096 samples): after 200 frames each camera's cache
    irradiance is far closer to a 256-spp reference than a 4-spp frame."""
    from oracle import nrc as N
    w, h = 240, 135
    p = T.noise_objects_params(seed=1)
    sc = T.generate_scene(p)
    cams = np.concatenate([T.generate_camera(p, i, w, h) for i in range(2)])
    ctx.upload(sc)
    ctx.svo_build(np.floor(cams[0]["camera"][:3]) - 1024.0, 11)
    sun = T.seeded_sun(1)
    refs, plains = [], []
    for c in range(2):
        acc = np.zeros((h, w, 3), np.float64)
        for f in range(64):
            acc += ctx.frame(cams[c:c + 1], w, h, T.gi_params(sun, spp=4, frame=1000 + f))[2]
        refs.append(acc / 64)
        plains.append(ctx.frame(cams[c:c + 1], w, h, T.gi_params(sun, spp=4, frame=7)))
    ctx.nrc_create(T.nrc_config(origin=(0, 0, 0), scale=1 / 512, train_samples=1 << 16, train_batch=1 << 12,
                                flags=T.NRC_FLAG_NO_FALLBACK | T.NRC_FLAG_JOINT_TRAINING))
    buf = Buffers(ctx, 2, w, h)
    try:
        say("joint training: 2 cameras, 200 frames")
        for _ in range(200):
            ctx.frame_batch(cams, w, h, T.gi_params(sun, spp=4), buf.vis, buf.rad, buf.irr, flags=T.FRAME_NEURAL)
        vis, _, irr = buf.read()
    finally:
        buf.free()
    assert ctx.nrc_stats()["adam_steps"] == 200 * 16
    for c in range(2):
        m = vis[c] != T.VISIBILITY_EMPTY
        e_plain = N.rel_mse(plains[c][2][m], refs[c][m])
        e_nrc = N.rel_mse(irr[c][m], refs[c][m])
        say(f"camera {c}: rel-MSE plain 4 spp {e_plain:.3f}, jointly trained cache {e_nrc:.3f}")
        assert e_nrc < 0.15 * e_plain


def test_tuning_rejects_out_of_range_values(ctx):
    """tgx_tuning_set refuses values outside each knob's documented range
    and leaves the tuning as it was."""
    before = ctx.tuning().copy()
    bad = [("trace_refill", 0), ("trace_refill", 65), ("trace_blocks_per_cu", 3), ("trace_sort", 4),
           ("trace_regroup", 9), ("trace_skip", 2), ("vis_setup_blocks", 32), ("vis_setup_blocks", 20000),
           ("trace_dda", 2), ("vis_tile", 12), ("vis_tile", 64), ("trace_cross", 65), ("trace_node", 0),
           ("trace_node", 5), ("trace_burst", 3), ("trace_burst", 34), ("trace_burst", 0), ("trace_chunk", 257),
           ("trace_small", 65)]
    for name, value in bad:
        with pytest.raises(T.TgxError) as e:
            ctx.set_tuning(**{name: value})
        assert e.value.code == 1, name
        assert ctx.tuning().tobytes() == before.tobytes(), name
    ctx.set_tuning(vis_tile=8, trace_node=4, trace_burst=32, trace_cross=64, trace_chunk=256)
    t = ctx.tuning()
    assert (t["vis_tile"][0], t["trace_node"][0], t["trace_burst"][0], t["trace_cross"][0],
            t["trace_chunk"][0]) == (8, 4, 32, 64, 256)
    ctx.set_tuning(**{k: int(before[k][0]) for k in before.dtype.names})
