"""Pins the CPU oracle of the one-bounce GI estimator (defined at
tgx_gi_params in include/tgx.h; tg has no GI of its own).

Closed-form scenes: a closed box (every bounce ray hits, no sunlight: zero
radiance), an open floor (every bounce ray escapes), a floor with a wall
lit from one side (light only on the lit side), plus exact linearity in
the sun radiance and the cosine lobe's first moment. The magnitude is fixed
by a floor beside a sunlit wall, against the float64 configuration factor
(Lambert's contour formula, itself checked against the parallel-rectangle
closed form, the infinite-wall limit 1/2 and cosine-weighted sampling).
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


def config_factor(p, n, verts) -> float:
    """Configuration factor from a differential area at p (unit normal n) to
    a planar polygon lying in its upper hemisphere, by Lambert's contour
    formula F = 1/(2 pi) sum_i gamma_i n . (R_i x R_i+1) / |R_i x R_i+1|,
    gamma_i the angle between consecutive vertex vectors. float64."""
    p, n = np.asarray(p, np.float64), np.asarray(n, np.float64)
    v = [np.asarray(q, np.float64) - p for q in verts]
    f = 0.0
    for i in range(len(v)):
        a, b = v[i], v[(i + 1) % len(v)]
        c = np.cross(a, b)
        nc = np.linalg.norm(c)
        if nc == 0.0:
            continue
        f += np.arctan2(nc, a @ b) * (n @ c) / nc
    return abs(f) / (2.0 * np.pi)


def test_config_factor_matches_closed_forms_and_sampling():
    # element below the corner of a parallel a x b rectangle at height c
    # (textbook closed form)
    a, b, c = 3.0, 5.0, 2.0
    A, B = a / c, b / c
    closed = (A / np.sqrt(1 + A * A) * np.arctan(B / np.sqrt(1 + A * A))
              + B / np.sqrt(1 + B * B) * np.arctan(A / np.sqrt(1 + B * B))) / (2 * np.pi)
    got = config_factor((0, 0, 0), (0, 0, 1), [(0, 0, c), (a, 0, c), (a, b, c), (0, b, c)])
    assert abs(got - closed) < 1e-12
    # a huge wall standing on the element's horizon takes half the cosine lobe
    big = 1e7
    half = config_factor((0, 0, 0), (0, 1, 0), [(1, 0, -big), (1, big, -big), (1, big, big), (1, 0, big)])
    assert abs(half - 0.5) < 1e-6
    # a finite perpendicular wall against cosine-weighted sampling (numpy)
    rng = np.random.default_rng(3)
    u1, u2 = rng.random(2_000_000), rng.random(2_000_000)
    r, phi = np.sqrt(u1), 2 * np.pi * u2
    d = np.stack([r * np.cos(phi), np.sqrt(1 - u1), r * np.sin(phi)], 1)  # cosine lobe about +y
    t = 4.0 / d[:, 0]
    hit = (d[:, 0] > 0) & (t * d[:, 1] <= 6.0) & (np.abs(t * d[:, 2] - 1.0) <= 3.0)
    f = config_factor((0, 0, 0), (0, 1, 0), [(4, 0, -2), (4, 6, -2), (4, 6, 4), (4, 0, 4)])
    assert abs(hit.mean() - f) < 4 * np.sqrt(f * (1 - f) / hit.size)


def test_irradiance_magnitude_beside_a_sunlit_wall():
    """Closed-form magnitude of the estimator: floor pixels beside a tall,
    wide wall whose -x face alone is sunlit. A cosine-weighted bounce ray
    from the floor's origin P (the hit point plus sqrt(3) along +y) sees
    only that face, so E(P) = F(P) a_wall (n_wall . s) L / pi with F the
    float64 configuration factor from P to the face above P's horizon, and
    the radiance is a_floor * E per channel. The oracle's mean over 16
    frames x 64 spp per pixel must agree within 3 sigma of the binomial
    sampling error; removing 1/pi, the wall albedo or the cosine changes it
    by far more, and removing the primary albedo fails the per-pixel
    product."""
    full = np.full(16, 0xFFFFFFFF, np.uint32)
    objs = [{"n": (4, 1, 4), "translation": (-8.0, -4.0, 0.0)},     # floor: x -24..8, y -8..0, z -16..16
            {"n": (1, 12, 12), "translation": (20.0, 40.0, 0.0)}]   # wall: x 16..24, y -8..88, z -48..48
    n_floor, n_wall = 16, 144
    sc = S.make_scene(objs, bits_per_cluster=[full] * (n_floor + n_wall), seed=9)
    sc.materials[:512 * n_floor] = 1
    sc.materials[512 * n_floor:] = 2
    a_floor = np.array([0.5, 0.7, 0.9], np.float32)
    a_wall = np.array([0.8, 0.6, 0.3], np.float32)
    sc.luts["albedo"][1], sc.luts["albedo"][2] = a_floor, a_wall
    cam = T.make_camera((-8.0, 40.0, 0.0), pitch=-1.5, yaw=0.0, roll=0.0, fov_y=0.5, aspect=1.0)
    sun = np.array([-0.6, 0.7, 0.3])
    w = h = 16
    spp, frames = 64, 16
    os_ = O.OracleScene(sc)
    vis, _ = O.visibility_culled(os_, cam, w, h)
    svo = O.OracleSvo(os_, (-64.0, -64.0, -64.0), 8)
    irr_sum = np.zeros((h, w, 3), np.float64)
    for f in range(frames):
        prm = T.gi_params(sun, spp=spp, frame=f)
        rad, irr, cnt = svo.gi(cam, w, h, vis, prm)
        assert cnt[0] == w * h
        np.testing.assert_array_equal(rad, irr * a_floor)  # L = a_p E, per channel, in fp32
        irr_sum += irr
    s32 = prm["sun_dir"][0].astype(np.float64)
    unit = a_wall.astype(np.float64) * (-s32[0]) * 1.0 / np.pi  # contribution of one lit sample
    expect = np.zeros(3)
    var = f_sum = 0.0
    for y in range(h):
        for x in range(w):
            ok, p, n, a = _surface(os_, cam, w, h, x, y, int(vis[y, x]))
            assert ok and n[1] == 1.0 and np.array_equal(a, a_floor)
            o = p.astype(np.float64) + np.sqrt(3.0) * n.astype(np.float64)
            F = config_factor(o, (0, 1, 0), [(16, o[1], -48), (16, 88, -48), (16, 88, 48), (16, o[1], 48)])
            expect += F * unit
            f_sum += F
            var += F * (1 - F) / (spp * frames)
    got = irr_sum.sum(axis=(0, 1)) / frames
    sigma = np.sqrt(var) * unit
    assert np.all(np.abs(got - expect) < 3 * sigma + 1e-4 * expect), (got, expect, sigma)
    assert 0.2 < f_sum / (w * h) < 0.5  # the wall takes a sizeable share of the lobe


def test_alg_bytes_surface_part_and_written_bytes_exact():
    """The surface part of tgxo_gi_alg_bytes recounted in numpy: every
    visibility word, and for each visible pixel the lines of its cluster
    pointer, cluster map entry, object record (96 B, up to two lines), voxel
    material and LUT entry."""
    p = T.noise_objects_params(seed=6, n_objects=30, max_extent=5, world=96.0, n_luts=3)
    sc = T.generate_scene(p)
    w, h = 64, 48
    cam = T.generate_camera(p, 0, w, h)
    os_ = O.OracleScene(sc)
    vis, _ = O.visibility_culled(os_, cam, w, h)
    svo = O.OracleSvo(os_, np.floor(cam[0]["camera"][:3]) - 64.0, 7)
    prm = T.gi_params(T.seeded_sun(6), spp=2)
    b = svo.gi_alg_bytes(cam, w, h, vis, prm)
    lines = {k: set() for k in ("cptr", "c2o", "obj", "mat", "lut")}
    for y in range(h):
        for x in range(w):
            word = int(vis[y, x])
            if not _surface(os_, cam, w, h, x, y, word)[0]:
                continue
            cptr, v = (word >> 9) & 0x7FFFFFFF, word & 511
            cidx = int(sc.cluster_pointers[cptr])
            obj = int(sc.cluster_idx_to_object[cidx])
            lines["cptr"].add(cptr * 4 // 64)
            lines["c2o"].add(cidx * 4 // 64)
            lines["obj"] |= {obj * 96 // 64, (obj * 96 + 95) // 64}
            lines["mat"].add((cidx * 512 + v) // 64)
            m = int(sc.materials[cidx * 512 + v])
            lines["lut"].add((int(sc.objects[obj]["lut_idx"]) * 256 + m) * 16 // 64)
    assert len(lines["cptr"]) > 3
    expect = 64 * (-(-8 * w * h // 64) + sum(len(s) for s in lines.values()))
    assert b["surface"] == expect
    assert b["lines"]["luts_surfaces"] == len(lines["lut"]) and b["lines"]["objects"] == len(lines["obj"])
    assert b["written"] == 12 * w * h
    assert svo.gi_alg_bytes(cam, w, h, vis, prm, with_irradiance=True)["written"] == 24 * w * h
    info = svo.info()
    n = b["lines"]
    assert 1 <= n["nodes"] <= -(-4 * int(info["n_nodes"]) // 64)
    assert 1 <= n["voxel_pool"] <= 64 * int(info["n_leaves"])
    assert n["leaf_data"] <= int(info["n_leaves"]) and n["leaf_materials"] <= 1024 * int(info["n_leaves"])
    assert b["trace"] == 64 * (n["nodes"] + n["leaf_data"] + n["voxel_pool"] + n["leaf_materials"] + n["luts_hits"])
    assert b["total"] == b["surface"] + b["trace"] + b["written"]


def test_alg_bytes_hit_shading_lines():
    """Traversal part: in the closed box no bounce hit is lit, so no leaf
    material or LUT line is read for hits while the voxel pool is; beside the
    sunlit wall every lit hit has the wall's one material, one LUT line."""
    sc = scene_from_grid(np.pad(np.zeros((28, 28, 28), bool), 2, constant_values=True), (0.0, 0.0, 0.0))
    cam = T.make_camera((0.5, 0.5, 0.5), pitch=0.2, yaw=0.7, roll=0.0, fov_y=1.4, aspect=1.0)
    os_ = O.OracleScene(sc)
    vis, _ = O.visibility_culled(os_, cam, 32, 32)
    b = O.OracleSvo(os_, (-64.0, -64.0, -64.0), 7).gi_alg_bytes(cam, 32, 32, vis, T.gi_params(spp=4))
    assert b["lines"]["leaf_materials"] == 0 and b["lines"]["luts_hits"] == 0
    assert b["lines"]["voxel_pool"] > 0 and b["lines"]["leaf_data"] > 0
    full = np.full(16, 0xFFFFFFFF, np.uint32)
    objs = [{"n": (4, 1, 4), "translation": (-8.0, -4.0, 0.0)}, {"n": (1, 12, 12), "translation": (20.0, 40.0, 0.0)}]
    sc = S.make_scene(objs, bits_per_cluster=[full] * 160, seed=9)
    sc.materials[:512 * 16] = 1
    sc.materials[512 * 16:] = 2
    cam = T.make_camera((-8.0, 40.0, 0.0), pitch=-1.5, yaw=0.0, roll=0.0, fov_y=0.5, aspect=1.0)
    os_ = O.OracleScene(sc)
    vis, _ = O.visibility_culled(os_, cam, 16, 16)
    b = O.OracleSvo(os_, (-64.0, -64.0, -64.0), 8).gi_alg_bytes(cam, 16, 16, vis,
                                                                T.gi_params((-0.6, 0.7, 0.3), spp=16))
    assert b["lines"]["luts_hits"] == 1 and b["lines"]["luts_surfaces"] == 1
    assert b["lines"]["leaf_materials"] > 0


def test_surface_count_equals_per_pixel_surfaces():
    """tgxo_gi_surface_count counts the pixels tgxo_gi_surface accepts: a
    word at depth exactly 1 (depth field all ones) is a buffer hit but no
    surface, and empty words are skipped."""
    p = T.noise_objects_params(seed=2, n_objects=20, max_extent=4, world=80.0)
    sc = T.generate_scene(p)
    cam = T.generate_camera(p, 0, 40, 30)
    os_ = O.OracleScene(sc)
    vis, _ = O.visibility_culled(os_, cam, 40, 30)
    hits = np.argwhere(vis != T.VISIBILITY_EMPTY)
    assert len(hits) > 50
    y0, x0 = hits[0]
    vis[y0, x0] = np.uint64((0xFFFFFF << 40) | (int(vis[y0, x0]) & ((1 << 40) - 1)))
    want = sum(_surface(os_, cam, 40, 30, x, y, int(vis[y, x]))[0]
               for y in range(30) for x in range(40) if vis[y, x] != T.VISIBILITY_EMPTY)
    assert want == len(hits) - 1
    assert O.gi_surface_count(os_, cam, 40, 30, vis, 1) == want
    assert O.gi_surface_count(os_, cam, 40, 30, vis, 3) == want
