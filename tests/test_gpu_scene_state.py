"""Scene replacement on the GPU: what a new scene leaves behind of the old.

- A scene uploaded without materials after one with materials shades,
  builds and traces with material 0 everywhere (the oracle on a copy with
  zero materials), not with the earlier scene's buffer, and reports
  has_materials = 0.
- A scene without LUTs is refused by the calls that read LUTs.
- Replacing the scene or editing objects makes the SVO stale: tracing calls
  return TGX_ERR_NOT_BUILT until tgx_svo_update, whose result equals a
  fresh build's.
- Scene replacement is refused while a frame graph exists."""
from __future__ import annotations

import ctypes
import os
import sys
from dataclasses import replace

import numpy as np
import pytest

import oracle as O
import tg_amd as T

pytestmark = pytest.mark.gpu

NOT_BUILT, CAPACITY, INVALID_ARG = 7, 2, 1


def say(msg: str) -> None:
    print(f"[gpu] {msg}", file=sys.stderr, flush=True)


def upload_raw(ctx, sc, materials: bool = True, luts: bool = True) -> None:
    d = sc.desc()
    if not materials:
        d.materials = None
    if not luts:
        d.luts, d.n_luts = None, 0
    ctx._scene_ref = sc
    ctx._check(ctx.lib.tgx_scene_upload(ctx.h, ctypes.byref(d)), "tgx_scene_upload")


def test_scene_without_materials_after_one_with_materials(ctx):
    small = T.generate_scene(T.noise_objects_params(seed=3, n_objects=20, world=128.0, n_luts=4))
    big = T.generate_scene(T.noise_objects_params(seed=4, n_objects=60, world=128.0, n_luts=4))
    assert big.n_clusters > small.n_clusters and small.materials.any()
    cam = T.generate_camera(T.noise_objects_params(seed=4, n_objects=60, world=128.0, n_luts=4), 0, 160, 90)
    say("upload with materials, then a larger scene without")
    ctx.upload(small)
    upload_raw(ctx, big, materials=False)
    assert int(ctx.scene_counts()["has_materials"]) == 0
    assert not ctx.scene_arrays().materials.any()
    zero = replace(big, materials=np.zeros_like(big.materials))
    osc = O.OracleScene(zero)
    vis = ctx.visibility(cam, 160, 90)[0]
    assert (vis != T.VISIBILITY_EMPTY).mean() > 0.05
    for view in (T.SHADE_LIT, T.SHADE_COLOR_LUT_INDEX, T.SHADE_COLOR):
        got = ctx.shade(cam, 160, 90, vis, T.shade_params(view, roughness=-1.0))
        ref = O.shade(osc, cam, 160, 90, vis, T.shade_params(view, roughness=-1.0))
        assert got.tobytes() == ref.tobytes(), f"view {view}"
    say("SVO build and GI frame of the material-less scene")
    mn = np.floor(cam[0]["camera"][:3]) - 128.0
    ctx.svo_build(mn, 8)
    ref_svo = O.OracleSvo(osc, mn, 8)
    ex = ctx.svo_export()
    re = ref_svo.export()
    assert np.array_equal(ex["materials"], re.materials)
    prm = T.gi_params(T.seeded_sun(2), spp=2)
    v, r, i = ctx.frame(cam, 160, 90, prm)
    rr, ri, _ = ref_svo.gi(cam, 160, 90, v, prm)
    assert r.tobytes() == rr.tobytes() and i.tobytes() == ri.tobytes() and r.any()


def test_scene_without_luts_is_refused_by_lut_readers(ctx):
    p = T.single_object_params(2)
    sc = T.generate_scene(p)
    cam = T.generate_camera(p, 0, 32, 32)
    upload_raw(ctx, sc, luts=False)
    assert int(ctx.scene_counts()["n_luts"]) == 0
    vis = ctx.visibility(cam, 32, 32)[0]  # visibility needs no LUT
    with pytest.raises(T.TgxError) as e:
        ctx.shade(cam, 32, 32, vis, T.shade_params())
    assert e.value.code == INVALID_ARG
    ctx.svo_build((-64.0, -64.0, -64.0), 7)
    with pytest.raises(T.TgxError) as e:
        ctx.frame(cam, 32, 32, T.gi_params(spp=1))
    assert e.value.code == INVALID_ARG


def test_replaced_scene_makes_the_svo_stale(ctx):
    pa = T.noise_objects_params(seed=5, n_objects=16, world=128.0)
    pb = T.noise_objects_params(seed=6, n_objects=16, world=128.0)
    a, b = T.generate_scene(pa), T.generate_scene(pb)
    cam = T.generate_camera(pb, 0, 96, 64)
    mn = np.floor(cam[0]["camera"][:3]) - 128.0
    prm = T.gi_params(T.seeded_sun(3), spp=2)
    o = np.repeat(cam[0]["camera"][None, :3], 4, 0)
    d = np.eye(3, dtype=np.float32)[[0, 1, 2, 0]]
    ctx.upload(a)
    ctx.svo_build(mn, 8)
    ctx.frame(cam, 96, 64, prm)
    say("upload of another scene: tracing calls refuse the old SVO")
    ctx.upload(b)
    for call in (lambda: ctx.frame(cam, 96, 64, prm), lambda: ctx.svo_trace(o, d),
                 lambda: ctx.svo_visibility(cam, 96, 64)):
        with pytest.raises(T.TgxError) as e:
            call()
        assert e.value.code == NOT_BUILT
    with pytest.raises(T.TgxError) as e:
        ctx.nrc_create(T.nrc_config())
        ctx.nrc_frame(cam, 96, 64, prm)
    assert e.value.code == NOT_BUILT
    ctx.svo_update()
    v1, r1, i1 = ctx.frame(cam, 96, 64, prm)
    ctx.svo_build(mn, 8)
    v2, r2, i2 = ctx.frame(cam, 96, 64, prm)
    assert v1.tobytes() == v2.tobytes() and r1.tobytes() == r2.tobytes() and r1.any()
    rr, _, _ = O.OracleSvo(O.OracleScene(b), mn, 8).gi(cam, 96, 64, v2, prm)
    assert r2.tobytes() == rr.tobytes()


def test_object_edits_make_the_svo_stale(ctx):
    rng = np.random.default_rng(11)
    luts = np.zeros(256, T.LUT_DTYPE)
    luts["albedo"] = rng.uniform(0, 1, (256, 3))
    ctx.scene_reserve(8, 64, luts)
    eye = np.eye(4, dtype=np.float32).reshape(-1)
    bits = rng.integers(0, 2**32, 16 * 8, dtype=np.uint32)
    ctx.object_create((2, 2, 2), (0.0, 0.0, 0.0), eye, bits=bits)
    cam = T.make_camera((0.0, 0.0, 40.0), 0.0, 0.0, 0.0, np.radians(60), 1.0)
    prm = T.gi_params(spp=1)
    ctx.svo_build((-64.0, -64.0, -64.0), 7)
    ctx.frame(cam, 32, 32, prm)
    ctx.object_create((1, 1, 1), (20.0, 0.0, 0.0), eye, bits=bits[:16])
    with pytest.raises(T.TgxError) as e:
        ctx.frame(cam, 32, 32, prm)
    assert e.value.code == NOT_BUILT
    ctx.svo_update()
    ctx.frame(cam, 32, 32, prm)


def test_scene_replacement_refused_while_a_graph_exists():
    p = T.noise_objects_params(seed=7, n_objects=12, world=128.0)
    sc = T.generate_scene(p)
    cam = T.generate_camera(p, 0, 64, 48)
    mn = np.floor(cam[0]["camera"][:3]) - 128.0
    prm = T.gi_params(spp=1)
    ctx = T.Context(int(os.environ.get("TGX_DEVICE", "0")))
    bufs = [ctx.alloc(n * 64 * 48) for n in (8, 12, 12)]
    try:
        ctx.upload(sc)
        ctx.svo_build(mn, 8)
        g = ctx.frame_graph(cam, 64, 48, prm, *bufs)
        try:
            for call in (lambda: ctx.upload(sc), lambda: ctx.scene_reserve(4, 16, sc.luts[:256]),
                         lambda: ctx.generate_scene_device(p)):
                with pytest.raises(T.TgxError) as e:
                    call()
                assert e.value.code == CAPACITY
            g.launch()
            assert g.stats()["primary_rays"] == 64 * 48
        finally:
            g.destroy()
        ctx.upload(sc)  # unlocked again
        ctx.svo_update()
        ctx.frame(cam, 64, 48, prm)
    finally:
        for q in bufs:
            ctx.free(q)
        ctx.close()
