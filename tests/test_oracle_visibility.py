"""Pins the CPU oracle of primary visibility (visibility.frag:22-208).

Each check is chosen so that a plausible transcription mistake (a dropped
term, a wrong sign, a transposed operand, the wrong DDA axis) fails it:
closed forms on axis-aligned rays, a float64 grid walk that sorts plane
crossings, numpy linear algebra, and the exhaustive min over all
(pixel, cluster) pairs.
"""
from __future__ import annotations

import numpy as np
import pytest

import oracle as O
import tg_amd as T
from tests import scenes as S

F32 = np.float32


def expected_depth24(distance: float, far: float) -> int:
    """(u64)(max(0, d/far) * 16777215) in fp32 (visibility.frag:156,198)."""
    d = F32(distance) / F32(far)
    d = d if d > 0 else F32(0.0)
    return int(F32(d * F32(16777215.0)))


def word(depth24: int, cptr: int, vox: int) -> int:
    return (depth24 << 40) | (cptr << 9) | vox


# --------------------------------------------------------------------------
# m4 helpers
# --------------------------------------------------------------------------

def test_m4_mul_matches_float64_product():
    rng = np.random.default_rng(1)
    for _ in range(50):
        a = rng.normal(size=16).astype(np.float32)
        b = rng.normal(size=16).astype(np.float32)
        got = O.m4_mul(a, b).reshape(4, 4).T  # column-major -> row-major
        ref = a.reshape(4, 4).T.astype(np.float64) @ b.reshape(4, 4).T.astype(np.float64)
        np.testing.assert_allclose(got, ref, rtol=1e-5, atol=1e-5)


def test_m4_inverse_matches_numpy():
    rng = np.random.default_rng(2)
    for _ in range(100):
        m = rng.normal(size=(4, 4)) + 4 * np.eye(4)
        col_major = m.T.reshape(-1).astype(np.float32)
        got = O.m4_inverse(col_major).reshape(4, 4).T
        ref = np.linalg.inv(m.astype(np.float32).astype(np.float64))
        np.testing.assert_allclose(got, ref, rtol=2e-5, atol=2e-5)


def test_m4_inverse_of_identity_is_exact():
    eye = S.identity_rotation()
    assert np.array_equal(O.m4_inverse(eye), eye)


def test_world_to_cluster_maps_cluster_corners():
    rng = np.random.default_rng(3)
    for trial in range(20):
        n = rng.integers(1, 6, size=3)
        rot = S.random_rotation(rng)
        t = rng.uniform(-300, 300, size=3).astype(np.float32)
        sc = S.make_scene([{"n": n, "translation": t, "rotation": rot}])
        obj = sc.objects[0]
        r3 = rot.reshape(4, 4).T[:3, :3].astype(np.float64)
        for rel in rng.integers(0, int(np.prod(n)), size=4):
            rx, ry, rz = rel % n[0], (rel // n[0]) % n[1], rel // (n[0] * n[1])
            m = O.world_to_cluster(obj, int(rel)).reshape(4, 4).T.astype(np.float64)
            for corner in np.ndindex(2, 2, 2):
                local = -4.0 * n + 8.0 * np.array([rx, ry, rz]) + 8.0 * np.array(corner)
                world = r3 @ local + t
                cs = m @ np.append(world, 1.0)
                np.testing.assert_allclose(cs[:3], 8.0 * np.array(corner), atol=2e-3)


# --------------------------------------------------------------------------
# Slab test (collide.inc:3-24)
# --------------------------------------------------------------------------

def test_slab_closed_forms():
    hit, e, x = O.intersect_ray_aabb((-5, 0.5, 0.5), (1, 0, 0), (0, 0, 0), (1, 1, 1))
    assert hit and e == 5.0 and x == 6.0
    hit, e, x = O.intersect_ray_aabb((0.5, 0.5, 0.5), (0, 0, -1), (0, 0, 0), (1, 1, 1))
    assert hit and e == -0.5 and x == 0.5
    # parallel to the x slabs, outside them: the reference's zero-direction
    # guard (collide.inc:5-11) leaves x unconstrained, so the test passes
    hit, e, x = O.intersect_ray_aabb((2, 0.5, -3), (0, 0, 1), (0, 0, 0), (1, 1, 1))
    assert hit and e == 3.0 and x == 4.0
    # box behind the ray: exit < 0
    hit, e, x = O.intersect_ray_aabb((3, 0.5, 0.5), (1, 0, 0), (0, 0, 0), (1, 1, 1))
    assert not hit and x == -2.0
    # diagonal entry through the corner: enter = 1 exactly
    hit, e, x = O.intersect_ray_aabb((-1, -1, -1), (1, 1, 1), (0, 0, 0), (1, 1, 1))
    assert hit and e == 1.0 and x == 2.0


# --------------------------------------------------------------------------
# Single-pair closed forms
# --------------------------------------------------------------------------

def _unit_cluster_scene(voxels, translation=(4, 4, 4)):
    return S.make_scene([{"n": (1, 1, 1), "translation": translation}], bits_per_cluster=[S.voxel_bits(voxels)])


@pytest.mark.parametrize("axis,sign", [(0, 1), (0, -1), (1, 1), (1, -1), (2, 1), (2, -1)])
def test_axis_ray_hits_first_solid_voxel_at_closed_form_depth(axis, sign):
    # Column of voxels along `axis` through lateral voxel (3, 5) (in the two
    # other axes); voxels at positions 2 and 6 along the axis are solid.
    lat = [i for i in range(3) if i != axis]
    vox = []
    for p in (2, 6):
        v = [0, 0, 0]
        v[axis] = p
        v[lat[0]], v[lat[1]] = 3, 5
        vox.append(tuple(v))
    sc = _unit_cluster_scene(vox)
    pos = [0.0, 0.0, 0.0]
    pos[lat[0]], pos[lat[1]] = 3.5, 5.5
    pos[axis] = -100.0 if sign > 0 else 108.0
    cam = S.axis_camera(pos, axis, sign)
    os_ = O.OracleScene(sc)
    got = O.vis_pair(os_, cam, 1, 1, 0, 0, 0)
    # first voxel along the ray: position 2 for +, 6 for -; the entered face
    # is at 2 (coming from below) or 7 (coming from above).
    first = 2 if sign > 0 else 6
    face = 2.0 if sign > 0 else 7.0
    dist = face - pos[axis] if sign > 0 else pos[axis] - face
    v = [0, 0, 0]
    v[axis] = first
    v[lat[0]], v[lat[1]] = 3, 5
    rel = 64 * v[2] + 8 * v[1] + v[0]
    assert got == word(expected_depth24(dist, 1000.0), 0, rel)


def test_camera_inside_solid_voxel_has_depth_zero():
    sc = _unit_cluster_scene([(3, 5, 2)])
    cam = S.axis_camera((3.5, 5.5, 2.5), 2, -1)
    got = O.vis_pair(O.OracleScene(sc), cam, 1, 1, 0, 0, 0)
    assert got == word(0, 0, 64 * 2 + 8 * 5 + 3)


def test_far_plane_rejects_distant_hits():
    sc = _unit_cluster_scene([(3, 5, 2)])
    cam = S.axis_camera((3.5, 5.5, 100.0), 2, -1, far=50.0)  # hit at distance 97 > far
    assert O.vis_pair(O.OracleScene(sc), cam, 1, 1, 0, 0, 0) == int(T.VISIBILITY_EMPTY)
    cam = S.axis_camera((3.5, 5.5, 100.0), 2, -1, far=97.0)  # d == 1 exactly: kept
    assert O.vis_pair(O.OracleScene(sc), cam, 1, 1, 0, 0, 0) == word(16777215, 0, 64 * 2 + 8 * 5 + 3)


def test_ray_parallel_to_slab_outside_cluster_is_a_miss():
    # dir = (0, 0, -1) exactly, origin x = 9.5 outside the cluster's [0, 8]:
    # the bare slab test would pass; the pair must not produce a hit.
    sc = S.make_scene([{"n": (1, 1, 1), "translation": (4, 4, 4)}],
                      bits_per_cluster=[np.full(16, 0xFFFFFFFF, np.uint32)])
    cam = S.axis_camera((9.5, 5.5, 100.0), 2, -1)
    assert O.vis_pair(O.OracleScene(sc), cam, 1, 1, 0, 0, 0) == int(T.VISIBILITY_EMPTY)
    cam = S.axis_camera((8.0, 5.5, 100.0), 2, -1)  # on the face: inside the closed slab
    assert O.vis_pair(O.OracleScene(sc), cam, 1, 1, 0, 0, 0) == word(expected_depth24(92.0, 1000.0), 0, 64 * 7 + 8 * 5 + 7)


def test_empty_and_full_clusters():
    empty = S.make_scene([{"n": (1, 1, 1), "translation": (4, 4, 4)}], bits_per_cluster=[np.zeros(16, np.uint32)])
    cam = S.axis_camera((3.5, 5.5, 100.0), 2, -1)
    assert O.vis_pair(O.OracleScene(empty), cam, 1, 1, 0, 0, 0) == int(T.VISIBILITY_EMPTY)
    full = S.make_scene([{"n": (1, 1, 1), "translation": (4, 4, 4)}],
                        bits_per_cluster=[np.full(16, 0xFFFFFFFF, np.uint32)])
    got = O.vis_pair(O.OracleScene(full), cam, 1, 1, 0, 0, 0)
    assert got == word(expected_depth24(92.0, 1000.0), 0, 64 * 7 + 8 * 5 + 3)


def test_cluster_pointer_field_and_object_offset():
    # Object of 2x1x1 clusters: the ray through the second cluster reports
    # pointer 1 (not the cluster index) and cluster-local voxel coordinates.
    b0 = S.voxel_bits([(1, 1, 1)])
    b1 = S.voxel_bits([(3, 5, 2)])
    sc = S.make_scene([{"n": (2, 1, 1), "translation": (8, 4, 4)}], bits_per_cluster=[b0, b1],
                      cluster_perm=np.array([1, 0], np.uint32))
    cam = S.axis_camera((8 + 3.5, 5.5, 100.0), 2, -1)
    os_ = O.OracleScene(sc)
    assert O.vis_pair(os_, cam, 1, 1, 0, 0, 0) == int(T.VISIBILITY_EMPTY)
    assert O.vis_pair(os_, cam, 1, 1, 0, 0, 1) == word(expected_depth24(97.0, 1000.0), 1, 64 * 2 + 8 * 5 + 3)


# --------------------------------------------------------------------------
# DDA against a float64 walk that sorts all plane crossings
# --------------------------------------------------------------------------

def _f32_normalize(d):
    d = np.asarray(d, np.float32)
    mag = np.sqrt(F32(d[0] * d[0] + d[1] * d[1]) + F32(d[2] * d[2]), dtype=np.float32)
    return np.array([d[0] / mag, d[1] / mag, d[2] / mag], np.float32)


def _crossing_walk(o, d, solid):
    """First solid voxel of the ray o + t d (t > 0) through [0,8]^3, found by
    sorting the crossings with every integer plane; None if it misses, or
    'ambiguous' if two crossings are too close to call."""
    o = o.astype(np.float64)
    d = d.astype(np.float64)
    with np.errstate(divide="ignore", invalid="ignore"):
        t0 = np.where(d != 0, (0 - o) / d, -np.inf)
        t1 = np.where(d != 0, (8 - o) / d, np.inf)
    tn = np.minimum(t0, t1).max()
    tf = np.maximum(t0, t1).min()
    if not (tf > 0 and tn <= tf):
        return None
    ts = []
    for a in range(3):
        if d[a] == 0:
            continue
        for p in range(0, 9):
            t = (p - o[a]) / d[a]
            if tn < t < tf:
                ts.append(t)
    ts = sorted(ts + [max(tn, 0.0), tf])
    if min(np.diff(ts)) < 1e-4:
        return "ambiguous"
    for a, b in zip(ts[:-1], ts[1:]):
        if b <= max(tn, 0.0):
            continue
        m = o + 0.5 * (max(a, 0.0) + b) * d
        v = np.floor(m).astype(int)
        if np.any(v < 0) or np.any(v > 7):
            continue
        if solid[v[2], v[1], v[0]]:
            return tuple(int(x) for x in v)
    return None


def test_dda_matches_sorted_crossing_walk():
    rng = np.random.default_rng(5)
    checked = 0
    for trial in range(400):
        solid = rng.random((8, 8, 8)) < 0.15
        vox = [(x, y, z) for z, y, x in zip(*np.nonzero(solid))]
        sc = _unit_cluster_scene(vox)
        target = rng.uniform(0.0, 8.0, size=3)
        start = rng.uniform(-30, 38, size=3)
        if trial % 4 == 0:
            start = rng.uniform(0.2, 7.8, size=3)  # origin inside the cluster
        d = (target - start).astype(np.float32)
        if np.linalg.norm(d) < 1e-3:
            continue
        cam = S.direction_camera(start.astype(np.float32), d)
        got = O.vis_pair(O.OracleScene(sc), cam, 1, 1, 0, 0, 0)
        dn = _f32_normalize(d)
        exp = _crossing_walk(start.astype(np.float32), dn, solid)
        if exp == "ambiguous":
            continue
        checked += 1
        if exp is None:
            assert got == int(T.VISIBILITY_EMPTY), trial
        else:
            assert got != int(T.VISIBILITY_EMPTY), trial
            x, y, z = exp
            assert (got & 511) == 64 * z + 8 * y + x, trial
    assert checked > 300


# --------------------------------------------------------------------------
# Whole frames: culled == exhaustive; permutation invariance
# --------------------------------------------------------------------------

def test_single_object_culled_equals_brute():
    p = T.single_object_params(1)
    sc = T.generate_scene(p)
    os_ = O.OracleScene(sc)
    for ci in range(3):
        cam = T.generate_camera(p, ci, 96, 80)
        brute = O.visibility_brute(os_, cam, 96, 80)
        culled, _ = O.visibility_culled(os_, cam, 96, 80)
        assert np.array_equal(brute, culled)
        assert 0.05 < (brute != T.VISIBILITY_EMPTY).mean() < 0.95


def test_noise_scene_culled_equals_brute_low_res():
    p = T.noise_objects_params(seed=3, n_objects=48, max_extent=8, world=128.0)
    sc = T.generate_scene(p)
    os_ = O.OracleScene(sc)
    cam = T.generate_camera(p, 0, 80, 45)
    brute = O.visibility_brute(os_, cam, 80, 45)
    culled, _ = O.visibility_culled(os_, cam, 80, 45)
    assert np.array_equal(brute, culled)
    assert (brute != T.VISIBILITY_EMPTY).mean() > 0.2


def test_camera_inside_and_touching_clusters_culled_equals_brute():
    rng = np.random.default_rng(11)
    objs = [{"n": (3, 2, 2), "translation": (0.0, 0.0, 0.0), "rotation": S.random_rotation(rng)},
            {"n": (1, 1, 1), "translation": (0.0, 0.0, 4.0)},
            {"n": (2, 2, 1), "translation": (6.0, -1.0, 0.5), "rotation": S.random_rotation(rng)}]
    sc = S.make_scene(objs, seed=4)
    os_ = O.OracleScene(sc)
    positions = [(0.1, 0.2, 0.3), (0.0, 0.0, 8.0), (4.0, 4.0, 4.0), (-4.0, 0.0, 0.0), (6.0, -1.0, 0.5)]
    for i, pos in enumerate(positions):
        cam = T.make_camera(pos, pitch=0.3 * i - 0.5, yaw=1.1 * i, roll=0.0, fov_y=1.2, aspect=1.25)
        brute = O.visibility_brute(os_, cam, 40, 32)
        culled, _ = O.visibility_culled(os_, cam, 40, 32)
        assert np.array_equal(brute, culled), pos


def test_axis_aligned_camera_with_zero_direction_components_culled_equals_brute():
    # Identity rotations and an unrotated camera at odd resolution: the
    # centre column and row have exactly zero direction components.
    objs = [{"n": (2, 2, 2), "translation": (float(x), float(y), -40.0)}
            for x, y in ((-20, 0), (0, 0), (21, 5), (0, -18), (-8, 9))]
    sc = S.make_scene(objs, seed=8)
    os_ = O.OracleScene(sc)
    cam = T.make_camera((0.0, 0.0, 0.0), pitch=0.0, yaw=0.0, roll=0.0, fov_y=1.0, aspect=1.0)
    brute = O.visibility_brute(os_, cam, 33, 33)
    culled, _ = O.visibility_culled(os_, cam, 33, 33)
    assert np.array_equal(brute, culled)
    assert (brute != T.VISIBILITY_EMPTY).mean() > 0.05


def test_visibility_is_invariant_under_cluster_index_permutation():
    p = T.noise_objects_params(seed=7, n_objects=24, max_extent=6, world=96.0)
    sc = T.generate_scene(p)
    cam = T.generate_camera(p, 0, 128, 72)
    a, _ = O.visibility_culled(O.OracleScene(sc), cam, 128, 72)
    b, _ = O.visibility_culled(O.OracleScene(S.permute_clusters(sc, 99)), cam, 128, 72)
    assert np.array_equal(a, b)


def test_row_step_computes_only_sampled_rows():
    p = T.single_object_params(2)
    sc = T.generate_scene(p)
    cam = T.generate_camera(p, 0, 64, 64)
    full, _ = O.visibility_culled(O.OracleScene(sc), cam, 64, 64)
    part, _ = O.visibility_culled(O.OracleScene(sc), cam, 64, 64, row_step=4)
    assert np.array_equal(full[::4], part[::4])


def test_alg_bytes_counts_output_and_touched_lines():
    p = T.single_object_params(1)
    sc = T.generate_scene(p)
    cam = T.generate_camera(p, 0, 256, 256)
    b = O.visibility_alg_bytes(O.OracleScene(sc), cam, 256, 256)
    # output words + at most every input line of a 64-cluster scene
    assert 8 * 256 * 256 < b <= 8 * 256 * 256 + 64 * (64 + 4 + 4 + 2)
