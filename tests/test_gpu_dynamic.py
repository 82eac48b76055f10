"""Dynamic scenes on the GPU (tgx_scene_reserve / tgx_object_create /
tgx_object_destroy) against the bookkeeping model of oracle/dynamic.py:
after every create and destroy the device arrays equal the model's, the
visibility of the edited scene equals the oracle's on the model arrays, and
an SVO rebuilt after the edits, in full or by tgx_svo_update, equals the
oracle's."""
from __future__ import annotations

import sys

import numpy as np
import pytest

import oracle as O
import tg_amd as T
from oracle.dynamic import DynamicScene
from tests import scenes as S

pytestmark = pytest.mark.gpu

OBJ_CAP, CL_CAP = 12, 400


def say(msg: str) -> None:
    print(f"[gpu] {msg}", file=sys.stderr, flush=True)


def random_luts(rng, n=3):
    l = np.zeros(256 * n, T.LUT_DTYPE)
    l["albedo"] = rng.random((256 * n, 3), dtype=np.float32)
    l["roughness"] = rng.random(256 * n, dtype=np.float32)
    return l


def assert_same_arrays(got: T.Scene, ref: T.Scene, what: str) -> None:
    for f in ("objects", "cluster_pointers", "cluster_idx_to_object", "bits", "materials", "luts"):
        a, b = getattr(got, f), getattr(ref, f)
        assert a.shape == b.shape and a.tobytes() == b.tobytes(), f"{what}: {f} differs"


def test_create_destroy_sequence_matches_model(ctx):
    rng = np.random.default_rng(21)
    luts = random_luts(rng)
    ctx.scene_reserve(OBJ_CAP, CL_CAP, luts)
    model = DynamicScene(OBJ_CAP, CL_CAP, luts)
    assert_same_arrays(ctx.scene_arrays(), model.scene(), "empty")
    cam = T.make_camera((0.0, 0.0, -70.0), pitch=0.0, yaw=0.0, roll=0.0, fov_y=1.2, aspect=1.5)
    vis = ctx.visibility(cam, 96, 64)[0]
    assert np.all(vis == T.VISIBILITY_EMPTY)
    alive = []
    for step in range(40):
        free_cl = len(model.free_clusters)
        if alive and (rng.random() < 0.35 or not model.free_objects or free_cl < 27):
            o = alive.pop(int(rng.integers(len(alive))))
            say(f"step {step}: destroy object {o}")
            ctx.object_destroy(o)
            model.destroy(o)
        else:
            n = tuple(int(v) for v in rng.integers(1, 4, 3))
            k = int(np.prod(n))
            if k > free_cl or not model.free_objects:
                continue
            t = rng.uniform(-30, 30, 3).astype(np.float32)
            rot = S.random_rotation(rng)
            lut = int(rng.integers(3))
            bits = np.packbits(rng.random(512 * k) < 0.3, bitorder="little").view(np.uint32)
            mats = rng.integers(0, 256, 512 * k, dtype=np.uint8)
            o = ctx.object_create(n, t, rot, lut, bits, mats)
            say(f"step {step}: create {n} -> object {o}")
            assert o == model.create(n, t, rot, lut, bits, mats)
            alive.append(o)
        model.check_invariants()
        assert_same_arrays(ctx.scene_arrays(), model.scene(), f"step {step}")
        if step % 8 == 7 and model.n_cluster_pointers:
            got = ctx.visibility(cam, 96, 64)[0]
            ref, _ = O.visibility_culled(O.OracleScene(model.scene()), cam, 96, 64)
            assert np.array_equal(got, ref), f"step {step}: visibility differs from the oracle"
    assert model.n_cluster_pointers > 0
    # a destroy right after a visibility call moves the pointers the
    # visibility pass listed: the next call must see the new layout
    assert len(alive) >= 2
    ctx.visibility(cam, 96, 64)
    o = alive.pop(0)
    say(f"destroy object {o} right after a visibility call")
    ctx.object_destroy(o)
    model.destroy(o)
    got = ctx.visibility(cam, 96, 64)[0]
    ref, _ = O.visibility_culled(O.OracleScene(model.scene()), cam, 96, 64)
    assert np.array_equal(got, ref), "visibility after a destroy differs from the oracle"
    assert model.n_cluster_pointers > 0
    mn = np.array([-64.0, -64.0, -128.0], np.float32)
    ctx.svo_build(mn, 8)
    ref = O.OracleSvo(O.OracleScene(model.scene()), mn, 8).export()
    got = ctx.svo_export()
    assert np.array_equal(got["nodes"], ref.nodes)
    assert np.array_equal(got["bits"], ref.bits) and np.array_equal(got["materials"], ref.materials)


def test_dynamic_scene_errors(ctx):
    rng = np.random.default_rng(2)
    ctx.scene_reserve(2, 10, random_luts(rng, 1))
    eye = S.identity_rotation()
    with pytest.raises(T.TgxError):
        ctx.svo_build((-64.0, -64.0, -64.0), 7)  # no clusters yet
    a = ctx.object_create((2, 2, 2), (0, 0, 0), eye)
    with pytest.raises(T.TgxError):
        ctx.object_create((2, 2, 1), (0, 0, 0), eye)  # 4 clusters needed, 2 free
    with pytest.raises(T.TgxError):
        ctx.object_create((1, 1, 1), (0, 0, 0), eye, lut_idx=1)  # one LUT
    b = ctx.object_create((1, 1, 2), (0, 0, 0), eye)
    with pytest.raises(T.TgxError):
        ctx.object_create((1, 1, 1), (0, 0, 0), eye)  # no free object index
    ctx.object_destroy(a)
    with pytest.raises(T.TgxError):
        ctx.object_destroy(a)  # not alive
    assert int(ctx.scene_arrays().objects[b]["first_cluster_pointer"]) == 0
    ctx.object_destroy(b)
    assert ctx.scene_counts()["n_cluster_pointers"] == 0


SVO_KEYS = ("nodes", "bits", "materials", "leaf_n", "list_offsets", "list")


def test_svo_update_equals_full_build(ctx):
    """tgx_svo_update after creates and destroys equals a full tgx_svo_build
    of the edited scene byte for byte (and the oracle's SVO), and reuses the
    leaves an edit leaves untouched."""
    rng = np.random.default_rng(33)
    luts = random_luts(rng)
    ctx.scene_reserve(OBJ_CAP, CL_CAP, luts)
    model = DynamicScene(OBJ_CAP, CL_CAP, luts)
    mn = np.array([-128.0, -128.0, -128.0], np.float32)

    def create():
        n = tuple(int(v) for v in rng.integers(1, 4, 3))
        k = int(np.prod(n))
        t = rng.uniform(-100, 100, 3).astype(np.float32)
        rot = S.random_rotation(rng)
        lut = int(rng.integers(3))
        bits = np.packbits(rng.random(512 * k) < 0.4, bitorder="little").view(np.uint32)
        mats = rng.integers(0, 256, 512 * k, dtype=np.uint8)
        o = ctx.object_create(n, t, rot, lut, bits, mats)
        assert o == model.create(n, t, rot, lut, bits, mats)
        return o

    alive = [create() for _ in range(8)]
    with pytest.raises(T.TgxError):
        ctx.svo_update()  # nothing built yet
    ctx.svo_build(mn, 8)
    total_reused = 0
    for step in range(6):
        if step % 2 == 0:
            o = alive.pop(int(rng.integers(len(alive))))
            say(f"svo update step {step}: destroy {o}")
            ctx.object_destroy(o)
            model.destroy(o)
        else:
            alive.append(create())
            say(f"svo update step {step}: create {alive[-1]}")
        reused, voxelized = ctx.svo_update()
        info = ctx.svo_info()
        assert reused + voxelized == int(info["n_leaves"])
        total_reused += reused
        upd = ctx.svo_export()
        ref = O.OracleSvo(O.OracleScene(model.scene()), mn, 8).export()
        for k in SVO_KEYS:
            assert np.array_equal(upd[k], getattr(ref, k)), f"step {step}: {k} differs from the oracle"
        if step in (1, 4):
            # a full build in between: the next update compares against it
            ctx.svo_build(mn, 8)
            full = ctx.svo_export()
            for k in SVO_KEYS:
                assert np.array_equal(upd[k], full[k]), f"step {step}: {k} differs from a full build"
    assert total_reused > 0
    # an update right after a build reuses every leaf
    reused, voxelized = ctx.svo_update()
    assert voxelized == 0 and reused == int(ctx.svo_info()["n_leaves"])


def test_svo_update_without_previous_dynamic_build(ctx):
    """tgx_svo_update on an uploaded scene samples every leaf (nothing to
    compare against) and equals tgx_svo_build; with a frame graph alive it is
    refused like a build."""
    p = T.noise_objects_params(seed=3)
    sc = T.generate_scene(p)
    ctx.upload(sc)
    mn = np.array([-128.0, -128.0, -128.0], np.float32)
    ctx.svo_build(mn, 9)
    full = ctx.svo_export()
    reused, voxelized = ctx.svo_update()
    assert reused == 0 and voxelized == int(ctx.svo_info()["n_leaves"]) > 0
    upd = ctx.svo_export()
    for k in SVO_KEYS:
        assert np.array_equal(upd[k], full[k]), k
    w, h = 64, 36
    cam = T.generate_camera(p, 0, w, h)
    prm = T.gi_params(T.seeded_sun(1), spp=1)
    bufs = [ctx.alloc(8 * w * h), ctx.alloc(12 * w * h)]
    try:
        g = ctx.frame_graph(cam, w, h, prm, bufs[0], bufs[1])
        try:
            with pytest.raises(T.TgxError):
                ctx.svo_update()
        finally:
            g.destroy()
        ctx.svo_update()
    finally:
        for b in bufs:
            ctx.free(b)
