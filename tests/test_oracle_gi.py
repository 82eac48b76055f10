"""Pins the CPU oracle of the one-bounce GI estimator (defined at
tgx_gi_params in include/tgx.h; tg has no GI of its own).

Closed-form scenes: a closed box (every bounce ray hits, no sunlight: zero
radiance), an open floor (every bounce ray escapes), a floor with a wall
lit from one side (light only on the lit side), plus exact linearity in
the sun radiance and the cosine lobe's first moment.
"""
from __future__ import annotations

import numpy as np

import oracle as O
import tg_amd as T
from tests import scenes as S


def shell_bits(n: int, thickness: int, hollow: bool = True):
    """Per-cluster bits of an n x n x n-cluster object that is a hollow
    shell (or a single bottom layer when hollow is False)."""
    side = 8 * n
    g = np.zeros((side, side, side), bool)  # [z, y, x]
    if hollow:
        g[:] = True
        g[thickness:-thickness, thickness:-thickness, thickness:-thickness] = False
    out = []
    for rel in range(n ** 3):
        rx, ry, rz = rel % n, (rel // n) % n, rel // (n * n)
        sub = g[8 * rz:8 * rz + 8, 8 * ry:8 * ry + 8, 8 * rx:8 * rx + 8]
        out.append(np.packbits(sub.reshape(-1), bitorder="little").view(np.uint32))
    return out


def scene_from_grid(g: np.ndarray, translation):
    """Unrotated object whose voxels [z, y, x] are g (sides multiples of 8)."""
    nz, ny, nx = (s // 8 for s in g.shape)
    bits = []
    for rel in range(nx * ny * nz):
        rx, ry, rz = rel % nx, (rel // nx) % ny, rel // (nx * ny)
        sub = g[8 * rz:8 * rz + 8, 8 * ry:8 * ry + 8, 8 * rx:8 * rx + 8]
        bits.append(np.packbits(sub.reshape(-1), bitorder="little").view(np.uint32))
    return S.make_scene([{"n": (nx, ny, nz), "translation": translation}], bits_per_cluster=bits, seed=5)


def run_gi(sc, cam, w, h, params, mn, depth):
    os_ = O.OracleScene(sc)
    vis, _ = O.visibility_culled(os_, cam, w, h)
    svo = O.OracleSvo(os_, mn, depth)
    rad, irr, cnt = svo.gi(cam, w, h, vis, params)
    return vis, rad, irr, cnt


def test_closed_box_gets_no_light():
    sc = scene_from_grid(np.pad(np.zeros((28, 28, 28), bool), 2, constant_values=True), (0.0, 0.0, 0.0))
    cam = T.make_camera((0.5, 0.5, 0.5), pitch=0.2, yaw=0.7, roll=0.0, fov_y=1.4, aspect=1.0)
    vis, rad, irr, cnt = run_gi(sc, cam, 32, 32, T.gi_params(spp=4), (-64.0, -64.0, -64.0), 7)
    assert np.all(vis != T.VISIBILITY_EMPTY)
    assert cnt[0] == 32 * 32 and cnt[1] == 4 * cnt[0] and cnt[2] == cnt[1]
    assert np.all(rad == 0) and np.all(irr == 0)


def test_open_floor_bounces_escape():
    g = np.zeros((16, 8, 64), bool)
    g[:, 0, :] = True  # one-voxel floor, y = 0 of the object
    sc = scene_from_grid(g, (0.0, 0.0, 0.0))
    cam = T.make_camera((0.0, 20.0, 0.0), pitch=-1.2, yaw=0.0, roll=0.0, fov_y=1.0, aspect=1.0)
    vis, rad, irr, cnt = run_gi(sc, cam, 24, 24, T.gi_params((0.0, 1.0, 0.0), spp=8), (-64.0, -64.0, -64.0), 7)
    assert cnt[0] > 100 and cnt[2] == 0 and np.all(rad == 0)


def test_wall_lights_only_its_sunny_side():
    g = np.zeros((64, 32, 64), bool)
    g[:, 0, :] = True     # floor
    g[:, :, 32] = True    # wall at object x = 32 (world x = 0)
    sc = scene_from_grid(g, (0.0, 16.0, 0.0))  # object spans x -32..32, y 0..32, z -32..32
    cam = T.make_camera((0.0, 60.0, 0.0), pitch=-1.5, yaw=0.0, roll=0.0, fov_y=1.3, aspect=1.0)
    sun = (-1.0, 1.0, 0.0)  # lights the wall's -x face
    w = h = 48
    vis, rad, irr, cnt = run_gi(sc, cam, w, h, T.gi_params(sun, spp=16), (-64.0, -64.0, -64.0), 7)
    os_ = O.OracleScene(sc)
    lit_side = dark_side = 0
    for y in range(h):
        for x in range(w):
            word = int(vis[y, x])
            if word == int(T.VISIBILITY_EMPTY):
                continue
            ok, p, n, _ = _surface(os_, cam, w, h, x, y, word)
            if not ok or n[1] < 0.5:
                continue  # floor pixels only
            if p[0] < -2.0:
                lit_side += 1
                assert rad[y, x].sum() >= 0
            elif p[0] > 2.0:
                dark_side += 1
                assert np.all(rad[y, x] == 0), (x, y)
    assert lit_side > 100 and dark_side > 100
    assert rad[:, : w // 2].sum() > 0.5


def _surface(os_, cam, w, h, x, y, word):
    p, n, a = (np.zeros(3, np.float32) for _ in range(3))
    ok = O.load().tgxo_gi_surface(os_.ref, O._p(cam), w, h, x, y, word, O._p(p), O._p(n), O._p(a))
    return bool(ok), p, n, a


def test_radiance_is_linear_in_sun_radiance():
    p = T.noise_objects_params(seed=6, n_objects=30, max_extent=5, world=96.0)
    sc = T.generate_scene(p)
    cam = T.generate_camera(p, 0, 64, 48)
    mn = np.floor(cam[0]["camera"][:3]) - 64.0
    _, r1, _, c1 = run_gi(sc, cam, 64, 48, T.gi_params(spp=4, sun_radiance=1.0), mn, 7)
    _, r2, _, c2 = run_gi(sc, cam, 64, 48, T.gi_params(spp=4, sun_radiance=2.0), mn, 7)
    assert np.array_equal(c1, c2) and r1.sum() > 0
    assert np.array_equal(r2, 2.0 * r1)


def test_surface_normals_are_axis_faces_of_the_object():
    p = T.noise_objects_params(seed=2, n_objects=20, max_extent=4, world=80.0)
    sc = T.generate_scene(p)
    cam = T.generate_camera(p, 0, 40, 30)
    os_ = O.OracleScene(sc)
    vis, _ = O.visibility_culled(os_, cam, 40, 30)
    checked = 0
    for y in range(30):
        for x in range(40):
            word = int(vis[y, x])
            if word == int(T.VISIBILITY_EMPTY):
                continue
            ok, pt, n, a = _surface(os_, cam, 40, 30, x, y, word)
            assert ok
            obj = sc.objects[sc.cluster_idx_to_object[sc.cluster_pointers[(word >> 9) & 0x7FFFFFFF]]]
            r = obj["rotation"].reshape(4, 4).T[:3, :3]
            local = r.T @ n.astype(np.float64)
            assert np.isclose(np.abs(local).max(), 1.0, atol=1e-5) and np.isclose(np.linalg.norm(n), 1, atol=1e-6)
            # the hit point lies on the primary ray, at the decoded depth
            d = (word >> 40) / 16777215.0 * cam[0]["far_plane"]
            assert abs(np.linalg.norm(pt - cam[0]["camera"][:3]) - d) < 0.05 + 1e-3 * d
            checked += 1
    assert checked > 200


def test_cosine_lobe_first_moment():
    n = np.array([0.0, 0.0, 1.0], np.float32)
    cosines = []
    for i in range(20000):
        d = np.zeros(3, np.float32)
        O.load().tgxo_gi_direction(O._p(n), O.load().tgxo_gi_seed(i, 0, 1, 0), O._p(d))
        cosines.append(float(d @ n))
    c = np.array(cosines)
    assert c.min() >= 0.0
    assert abs(c.mean() - 2.0 / 3.0) < 0.01          # E[cos] = 2/3 for a cosine lobe
    assert abs((c ** 2).mean() - 0.5) < 0.01         # E[cos^2] = 1/2
