"""GPU parity of primary visibility: libtgx.so (gfx950) against the CPU
oracle, bit-exact on every pixel, through the C ABI."""
from __future__ import annotations

import sys

import numpy as np
import pytest

import oracle as O
import tg_amd as T
from tests import scenes as S

pytestmark = pytest.mark.gpu


def say(msg: str) -> None:
    print(f"[gpu] {msg}", file=sys.stderr, flush=True)


def assert_same(gpu: np.ndarray, ref: np.ndarray, what: str) -> None:
    bad = np.argwhere(gpu != ref)
    assert bad.size == 0, f"{what}: {len(bad)} pixels differ, first {bad[:5].tolist()} " \
                          f"gpu={[hex(int(gpu[tuple(b)])) for b in bad[:3]]} ref={[hex(int(ref[tuple(b)])) for b in bad[:3]]}"


def test_single_object_256_bit_exact(ctx):
    p = T.single_object_params(1)
    sc = T.generate_scene(p)
    ctx.upload(sc)
    os_ = O.OracleScene(sc)
    for ci in range(3):
        say(f"single object 256x256 camera {ci}")
        cam = T.generate_camera(p, ci, 256, 256)
        gpu = ctx.visibility(cam, 256, 256)[0]
        ref = O.visibility_brute(os_, cam, 256, 256)
        assert_same(gpu, ref, f"camera {ci}")
        assert (ref != T.VISIBILITY_EMPTY).mean() > 0.1


@pytest.fixture(scope="module")
def noise_scene():
    p = T.noise_objects_params(seed=1)
    return p, T.generate_scene(p)


def test_noise_objects_1080p_bit_exact(ctx, noise_scene):
    p, sc = noise_scene
    ctx.upload(sc)
    os_ = O.OracleScene(sc)
    for ci in range(2):
        say(f"512 noise objects 1920x1080 camera {ci}")
        cam = T.generate_camera(p, ci, 1920, 1080)
        gpu = ctx.visibility(cam, 1920, 1080)[0]
        ref, _ = O.visibility_culled(os_, cam, 1920, 1080)
        assert_same(gpu, ref, f"camera {ci}")
        assert (ref != T.VISIBILITY_EMPTY).mean() > 0.3


def test_multi_camera_batch_equals_single_cameras(ctx, noise_scene):
    p, sc = noise_scene
    ctx.upload(sc)
    cams = np.concatenate([T.generate_camera(p, i, 640, 360) for i in range(5)])
    say("5-camera batch 640x360")
    batch = ctx.visibility(cams, 640, 360)
    os_ = O.OracleScene(sc)
    for i in range(5):
        ref, _ = O.visibility_culled(os_, cams[i:i + 1], 640, 360)
        assert_same(batch[i], ref, f"camera {i}")


def test_odd_resolution_and_permuted_clusters(ctx, noise_scene):
    p, sc = noise_scene
    perm = S.permute_clusters(sc, 5)
    cam = T.generate_camera(p, 3, 333, 197)
    ref, _ = O.visibility_culled(O.OracleScene(sc), cam, 333, 197)
    say("333x197, identity cluster map")
    ctx.upload(sc)
    assert_same(ctx.visibility(cam, 333, 197)[0], ref, "identity map")
    say("333x197, permuted cluster map")
    ctx.upload(perm)
    assert_same(ctx.visibility(cam, 333, 197)[0], ref, "permuted map")


def test_tile_list_overflow_stays_exact(ctx, noise_scene):
    p, sc = noise_scene
    ctx.upload(sc)
    cam = T.generate_camera(p, 1, 480, 270)
    ref, _ = O.visibility_culled(O.OracleScene(sc), cam, 480, 270)
    try:
        for cap in (1, 3000):
            say(f"480x270 with tile-list capacity {cap}")
            ctx.set_list_capacity(cap)
            assert_same(ctx.visibility(cam, 480, 270)[0], ref, f"capacity {cap}")
            assert ctx.last_stats().tile_entries > cap
    finally:
        ctx.set_list_capacity(0)
    assert_same(ctx.visibility(cam, 480, 270)[0], ref, "automatic capacity")


def test_camera_inside_and_near_clusters(ctx):
    rng = np.random.default_rng(11)
    objs = [{"n": (3, 2, 2), "translation": (0.0, 0.0, 0.0), "rotation": S.random_rotation(rng)},
            {"n": (1, 1, 1), "translation": (0.0, 0.0, 4.0)},
            {"n": (2, 2, 1), "translation": (6.0, -1.0, 0.5), "rotation": S.random_rotation(rng)}]
    sc = S.make_scene(objs, seed=4)
    ctx.upload(sc)
    os_ = O.OracleScene(sc)
    for i, pos in enumerate([(0.1, 0.2, 0.3), (0.0, 0.0, 8.0), (4.0, 4.0, 4.0), (-4.0, 0.0, 0.0)]):
        say(f"near-camera case {i} at {pos}")
        cam = T.make_camera(pos, pitch=0.3 * i - 0.5, yaw=1.1 * i, roll=0.0, fov_y=1.2, aspect=1.25)
        assert_same(ctx.visibility(cam, 40, 32)[0], O.visibility_brute(os_, cam, 40, 32), f"case {i}")


def test_zero_direction_components(ctx):
    objs = [{"n": (2, 2, 2), "translation": (float(x), float(y), -40.0)}
            for x, y in ((-20, 0), (0, 0), (21, 5), (0, -18), (-8, 9))]
    sc = S.make_scene(objs, seed=8)
    ctx.upload(sc)
    cam = T.make_camera((0.0, 0.0, 0.0), pitch=0.0, yaw=0.0, roll=0.0, fov_y=1.0, aspect=1.0)
    say("axis-aligned camera 33x33")
    assert_same(ctx.visibility(cam, 33, 33)[0], O.visibility_brute(O.OracleScene(sc), cam, 33, 33), "axis camera")


def test_device_output_and_clear(ctx):
    p = T.single_object_params(4)
    sc = T.generate_scene(p)
    ctx.upload(sc)
    cam = T.generate_camera(p, 0, 64, 48)
    dev = ctx.alloc(64 * 48 * 8)
    try:
        say("device-pointer output 64x48")
        ctx.visibility_device(cam, 64, 48, dev)
        ctx.synchronize()
        out = np.empty((48, 64), np.uint64)
        ctx.d2h(out, dev)
        assert_same(out, O.visibility_brute(O.OracleScene(sc), cam, 64, 48), "device output")
        rc = ctx.lib.tgx_visibility_clear_device(ctx.h, dev, 64 * 48)
        assert rc == 0
        ctx.d2h(out, dev)
        assert np.all(out == T.VISIBILITY_EMPTY)
    finally:
        ctx.free(dev)


def test_error_paths(ctx):
    sc = T.generate_scene(T.single_object_params(1))
    bad = T.Scene(sc.objects, sc.cluster_pointers.copy(), sc.cluster_idx_to_object, sc.bits, sc.materials, sc.luts)
    bad.cluster_pointers[3] = 64  # out of range
    with pytest.raises(T.TgxError) as e:
        ctx.upload(bad)
    assert e.value.code == 1
    cam = T.generate_camera(T.single_object_params(1), 0, 8, 8)
    with pytest.raises(T.TgxError) as e:
        ctx.visibility(cam, 0, 8)
    assert e.value.code == 1
    fresh = T.Context(0)
    try:
        with pytest.raises(T.TgxError) as e:
            fresh.visibility(cam, 8, 8)
        assert e.value.code == 4
    finally:
        fresh.close()


def test_object_cull_at_frustum_edges(ctx):
    """Objects astride each screen edge (inside, on and just beyond it), behind
    the camera and around the near plane: the per-object cull must drop only
    objects none of whose clusters can cover a pixel, so the frame equals the
    oracle's brute-force visibility word for word."""
    rng = np.random.default_rng(23)
    cam = T.make_camera((3.0, -2.0, 5.0), pitch=0.2, yaw=0.7, roll=0.1, fov_y=1.0, aspect=1.25, far=400.0)
    pos = cam["camera"][0, :3].astype(np.float64)
    c = {k: cam[k][0, :3].astype(np.float64) for k in ("ray_bl", "ray_br", "ray_tr", "ray_tl")}
    fwd = sum(c.values()) / 4.0
    fwd /= np.linalg.norm(fwd)
    edges = [(c["ray_bl"], c["ray_tl"], c["ray_bl"] - c["ray_br"]),   # left
             (c["ray_br"], c["ray_tr"], c["ray_br"] - c["ray_bl"]),   # right
             (c["ray_bl"], c["ray_br"], c["ray_bl"] - c["ray_tl"]),   # bottom
             (c["ray_tl"], c["ray_tr"], c["ray_tl"] - c["ray_bl"])]   # top
    objs = []
    for a, b, out in edges:
        out = out / np.linalg.norm(out)
        for d in (40.0, 90.0, 200.0):
            for s in (0.0, 0.5, 1.0, 1.08, 1.4):
                n = tuple(int(v) for v in rng.integers(1, 3, 3))
                h = 4.0 * np.sqrt(3.0) * max(n)
                ray = a + rng.uniform(0.1, 0.9) * (b - a)
                centre = pos + d * ray / np.linalg.norm(ray) + s * h * out
                objs.append({"n": n, "translation": tuple(centre), "rotation": S.random_rotation(rng)})
    near = []
    for d in (-30.0, -9.0, -2.0, 0.5, 1.5, 6.0):
        lat = rng.normal(size=3)
        near.append({"n": (1, 1, 2), "translation": tuple(pos + d * fwd + (2.0 + abs(d)) * lat),
                     "rotation": S.random_rotation(rng)})
    w, h_ = 128, 96
    for name, group, lo, hi in (("frustum edges", objs, 0.02, 0.9), ("behind and near", near, 0.0, 1.0)):
        sc = S.make_scene(group, seed=12)
        ctx.upload(sc)
        say(f"object cull, {name}: {len(group)} objects, {w}x{h_}")
        got = ctx.visibility(cam, w, h_)[0]
        ref = O.visibility_brute(O.OracleScene(sc), cam, w, h_)
        assert_same(got, ref, name)
        assert lo <= (ref != T.VISIBILITY_EMPTY).mean() <= hi, name


def test_object_cull_at_the_far_plane(ctx):
    """Objects around the far plane (nearest voxel just inside, on and just
    beyond it, and whole objects beyond it): the per-object far cull drops
    only objects that cannot hold a word with depth <= 1, so the frame equals
    the oracle's brute-force visibility word for word."""
    rng = np.random.default_rng(31)
    far = 60.0
    cam = T.make_camera((1.0, 2.0, -3.0), pitch=-0.1, yaw=0.4, roll=0.0, fov_y=1.1, aspect=1.5, far=far)
    pos = cam["camera"][0, :3].astype(np.float64)
    c = [cam[k][0, :3].astype(np.float64) for k in ("ray_bl", "ray_br", "ray_tr", "ray_tl")]
    objs = []
    for k in range(40):
        n = tuple(int(v) for v in rng.integers(1, 4, 3))
        ray = c[0] + rng.uniform(0.1, 0.9) * (c[1] - c[0]) + rng.uniform(0.1, 0.9) * (c[3] - c[0])
        ray /= np.linalg.norm(ray)
        half = 4.0 * np.array(n, np.float64)
        # the centre's distance puts the object's near side around the far plane
        d = far + rng.uniform(-1.0, 1.0) * float(np.linalg.norm(half)) + rng.uniform(-3.0, 30.0)
        objs.append({"n": n, "translation": tuple(pos + d * ray), "rotation": S.random_rotation(rng)})
    sc = S.make_scene(objs, seed=13)
    ctx.upload(sc)
    w, h = 120, 80
    say(f"object cull at the far plane: {len(objs)} objects, {w}x{h}")
    got = ctx.visibility(cam, w, h)[0]
    ref = O.visibility_brute(O.OracleScene(sc), cam, w, h)
    assert_same(got, ref, "far plane")
    hits = ref[ref != T.VISIBILITY_EMPTY]
    assert hits.size > 50
    # some hits lie in the last percent of the depth range
    assert ((hits >> np.uint64(40)) > np.uint64(int(0.97 * 16777215))).any()


@pytest.mark.parametrize("tile", [8, 32])
def test_tile_edges_bit_exact(ctx, noise_scene, tile):
    """Tiles of 8x8 and 32x32 pixels (one workgroup of 64 or 1024 lanes per
    tile; the global-list threshold and the first tile-list capacity follow
    the tile edge) give the oracle's words: 1080p, an odd resolution, a
    3-camera batch, a forced tile-list overflow and the allocation-free
    entry point."""
    p, sc = noise_scene
    ctx.upload(sc)
    os_ = O.OracleScene(sc)
    try:
        ctx.set_tuning(vis_tile=tile)
        for w, h, ci in ((1920, 1080, 0), (333, 197, 3)):
            say(f"tile {tile}: {w}x{h} camera {ci}")
            cam = T.generate_camera(p, ci, w, h)
            ref, _ = O.visibility_culled(os_, cam, w, h)
            assert_same(ctx.visibility(cam, w, h)[0], ref, f"tile {tile} {w}x{h}")
            assert ctx.last_stats().tile_size == tile
        cams = np.concatenate([T.generate_camera(p, i, 640, 360) for i in range(3)])
        refs = [O.visibility_culled(os_, cams[i:i + 1], 640, 360)[0] for i in range(3)]
        say(f"tile {tile}: 3-camera batch 640x360")
        batch = ctx.visibility(cams, 640, 360)
        for i in range(3):
            assert_same(batch[i], refs[i], f"tile {tile} camera {i}")
        try:
            say(f"tile {tile}: 3-camera batch with tile-list capacity 500")
            ctx.set_list_capacity(500)
            batch = ctx.visibility(cams, 640, 360)
            assert ctx.last_stats().tile_entries > 500
            for i in range(3):
                assert_same(batch[i], refs[i], f"tile {tile} overflow camera {i}")
        finally:
            ctx.set_list_capacity(0)
        dev = ctx.alloc(3 * 640 * 360 * 8)
        try:
            say(f"tile {tile}: allocation-free entry point, 3 cameras 640x360")
            ctx.visibility_device_noalloc(cams, 640, 360, dev)
            ctx.synchronize()
            out = np.empty((3, 360, 640), np.uint64)
            ctx.d2h(out, dev)
            for i in range(3):
                assert_same(out[i], refs[i], f"tile {tile} no-alloc camera {i}")
        finally:
            ctx.free(dev)
    finally:
        ctx.set_tuning(vis_tile=16)
