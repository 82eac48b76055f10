"""GPU shading pass (tgx_shade_device) against the oracle, bit for bit, in
every view: the decode, hash colours, normals and GGX lighting use the same
fp32 operations in the same order on both sides."""
from __future__ import annotations

import sys

import numpy as np
import pytest

import oracle as O
import tg_amd as T

pytestmark = pytest.mark.gpu

VIEWS = list(range(10))


def say(msg: str) -> None:
    print(f"[gpu] {msg}", file=sys.stderr, flush=True)


@pytest.mark.parametrize("scene,w,h", [("single", 256, 256), ("noise", 480, 270)])
def test_shade_views_bit_exact(ctx, scene, w, h):
    p = T.single_object_params(1) if scene == "single" else T.noise_objects_params(seed=1, n_luts=4)
    sc = T.generate_scene(p)
    cam = T.generate_camera(p, 0, w, h)
    ctx.upload(sc)
    vis = ctx.visibility(cam, w, h)[0]
    os_ = O.OracleScene(sc)
    assert (vis != T.VISIBILITY_EMPTY).mean() > 0.05
    for view in VIEWS:
        say(f"shade {scene} {w}x{h} view {view}")
        prm = T.shade_params(view, light_dir=(0.3, 0.8, -0.4), roughness=-1.0 if view == 0 else 0.8)
        got = ctx.shade(cam, w, h, vis, prm)
        ref = O.shade(os_, cam, w, h, vis, prm)
        assert got.tobytes() == ref.tobytes(), f"view {view}: {int((got != ref).any(-1).sum())} pixels differ"


def test_shade_rejects_bad_arguments(ctx):
    p = T.single_object_params(1)
    ctx.upload(T.generate_scene(p))
    cam = T.generate_camera(p, 0, 8, 8)
    vis = np.full((8, 8), T.VISIBILITY_EMPTY, np.uint64)
    with pytest.raises(T.TgxError):
        ctx.shade(cam, 8, 8, vis, T.shade_params(10))
    with pytest.raises(T.TgxError):
        ctx.shade(cam, 8, 8, vis, T.shade_params(light_dir=(0.0, 0.0, 0.0)))
    out = ctx.shade(cam, 8, 8, vis, T.shade_params())
    assert np.all(out == np.array([1.0, 0.0, 1.0, 1.0], np.float32))


def lerp_dir_f32(cam, x, y, w, h):
    """tg_lerp_corner_ray_directions_nn in fp32, the kernels' operation order."""
    f = np.float32
    fx = (f(x) + f(0.5)) / f(w)
    fy = f(1.0) - (f(y) + f(0.5)) / f(h)
    omy, omx = f(1.0) - fy, f(1.0) - fx
    c = cam[0]
    left = c["ray_bl"][:3] * omy + c["ray_tl"][:3] * fy
    right = c["ray_br"][:3] * omy + c["ray_tr"][:3] * fy
    return left * omx + right * fx


def test_pick_resolves_pointer_index_object_and_position(ctx):
    p = T.noise_objects_params(seed=2)
    sc = T.generate_scene(p)
    w, h = 480, 270
    cam = T.generate_camera(p, 0, w, h)
    ctx.upload(sc)
    vis_d = ctx.alloc(8 * w * h)
    try:
        ctx.visibility_device(cam, w, h, vis_d)
        vis = np.empty((h, w), np.uint64)
        ctx.d2h(vis, vis_d)
        rng = np.random.default_rng(4)
        xy = np.stack([rng.integers(0, w, 2000), rng.integers(0, h, 2000)], 1).astype(np.uint32)
        say("pick 2000 pixels")
        got = ctx.pick(vis_d, w, h, xy, cam)
        for (x, y), r in zip(xy, got):
            word = int(vis[y, x])
            hit, depth, cptr, vox = T.decode(word)
            assert bool(r["hit"]) == hit and np.float32(r["depth"]) == np.float32(depth)
            if not hit:
                assert r["cluster_pointer"] == r["cluster_idx"] == r["object_idx"] == r["voxel"] == 0xFFFFFFFF
                continue
            cidx = int(sc.cluster_pointers[cptr])
            assert (r["cluster_pointer"], r["voxel"], r["cluster_idx"]) == (cptr, vox, cidx)
            assert r["object_idx"] == sc.cluster_idx_to_object[cidx]
            t = np.float32(depth) * cam[0]["far_plane"]
            pos = cam[0]["camera"][:3] + t * lerp_dir_f32(cam, int(x), int(y), w, h)
            assert r["position"].tobytes() == pos.astype(np.float32).tobytes()
        assert got["hit"].sum() > 100
        no_cam = ctx.pick(vis_d, w, h, xy[:10])
        assert np.all(no_cam["position"] == 0)
        with pytest.raises(T.TgxError):
            ctx.pick(vis_d, w, h, [[w, 0]])
    finally:
        ctx.free(vis_d)


def test_lut_set_changes_one_entry(ctx):
    """tgx_lut_set (tg_raytracer_color_lut_set) replaces one entry of the
    device LUTs; the colour view then shows it for the voxels that use it."""
    p = T.single_object_params(1)
    sc = T.generate_scene(p)
    ctx.upload(sc)
    cam = T.generate_camera(p, 0, 96, 64)
    vis = ctx.visibility(cam, 96, 64)[0]
    before = ctx.shade(cam, 96, 64, vis, T.shade_params(T.SHADE_COLOR))
    hit = vis != T.VISIBILITY_EMPTY
    _, _, cptr, vox = T.decode(int(vis[hit][0]))
    cidx = int(sc.cluster_pointers[cptr])
    mat = int(sc.materials[512 * cidx + vox])
    ctx.lut_set(0, mat, (0.25, 0.5, 0.75), 0.125)
    luts = ctx.scene_arrays().luts
    ref = sc.luts.copy()
    ref[mat]["albedo"] = (0.25, 0.5, 0.75)
    ref[mat]["roughness"] = 0.125
    assert luts.tobytes() == ref.tobytes()
    after = ctx.shade(cam, 96, 64, vis, T.shade_params(T.SHADE_COLOR))
    changed = (after != before).any(-1)
    assert changed.any() and changed[hit].sum() == changed.sum()
    with pytest.raises(T.TgxError):
        ctx.lut_set(1, 0, (0, 0, 0), 0)  # one LUT
    with pytest.raises(T.TgxError):
        ctx.lut_set(0, 256, (0, 0, 0), 0)
