"""Pins the CPU oracle of the shading pass (shading.frag:114-337).

Checks that a plausible transcription mistake fails: the debug views
against numpy restatements of the decode and of tg_hash_u32 (bit for bit),
face normals against float64 ray/box geometry (axis rays and rotated
objects), and the lit views against an independent float64 Cook-Torrance
evaluation (GGX distribution, Schlick-GGX geometry with k = (r + 1)^2 / 8,
tg's Fresnel variant f = (1 - r) + r (1 - n.v)^5), with the closed form
"light below the surface gives ambient only" exactly.
"""
from __future__ import annotations

import numpy as np
import pytest

import oracle as O
import tg_amd as T
from tests import scenes as S

F32 = np.float32
U32 = np.uint32


def murmur_fmix32(v: int) -> int:
    v &= 0xFFFFFFFF
    v ^= v >> 16
    v = (v * 0x85EBCA6B) & 0xFFFFFFFF
    v ^= v >> 13
    v = (v * 0xC2B2AE35) & 0xFFFFFFFF
    v ^= v >> 16
    return v


def hash_rgb(v: int) -> np.ndarray:
    h0 = murmur_fmix32(v)
    h1 = murmur_fmix32(h0)
    h2 = murmur_fmix32(h1)
    # f32(h) / 4294967295.0 in fp32: the literal rounds to 2^32
    return np.array([F32(h0) / F32(4294967295.0), F32(h1) / F32(4294967295.0), F32(h2) / F32(4294967295.0), 1.0],
                    np.float32)


def noise_setup(w=96, h=64):
    p = T.noise_objects_params(seed=3, n_objects=40, max_extent=5, world=96.0, n_luts=4)
    sc = T.generate_scene(p)
    cam = T.generate_camera(p, 0, w, h)
    os_ = O.OracleScene(sc)
    vis, _ = O.visibility_culled(os_, cam, w, h)
    return sc, os_, cam, vis


def decode_all(sc, vis):
    """(mask, object, cluster, voxel, material, depth) per pixel, numpy."""
    m = vis != T.VISIBILITY_EMPTY
    w = vis[m]
    cptr = ((w >> np.uint64(9)) & np.uint64(0x7FFFFFFF)).astype(np.int64)
    vox = (w & np.uint64(511)).astype(np.int64)
    cidx = sc.cluster_pointers[cptr].astype(np.int64)
    obj = sc.cluster_idx_to_object[cidx].astype(np.int64)
    mat = sc.materials[cidx * 512 + vox].astype(np.int64)
    depth = (w >> np.uint64(40)).astype(np.float32) / F32(16777215.0)
    return m, obj, cidx, vox, mat, depth


def test_miss_pixels_get_the_miss_colour():
    sc, os_, cam, vis = noise_setup()
    prm = T.shade_params(miss=(0.25, 0.5, 0.75, 1.0))
    out = O.shade(os_, cam, 96, 64, vis, prm)
    miss = vis == T.VISIBILITY_EMPTY
    assert miss.any() and (~miss).any()
    assert np.all(out[miss] == np.array([0.25, 0.5, 0.75, 1.0], np.float32))


@pytest.mark.parametrize("view,field", [(T.SHADE_OBJECT_INDEX, 1), (T.SHADE_CLUSTER_INDEX, 2),
                                        (T.SHADE_BLOCKS, 2), (T.SHADE_VOXEL_INDEX, 3),
                                        (T.SHADE_COLOR_LUT_INDEX, 4)])
def test_hash_views_match_the_murmur_finaliser(view, field):
    sc, os_, cam, vis = noise_setup()
    out = O.shade(os_, cam, 96, 64, vis, T.shade_params(view))
    dec = decode_all(sc, vis)
    got = out[dec[0]]
    keys = dec[field]
    for i in range(0, len(keys), 7):
        assert got[i].tobytes() == hash_rgb(int(keys[i])).tobytes()


def test_depth_and_colour_views():
    sc, os_, cam, vis = noise_setup()
    m, obj, cidx, vox, mat, depth = decode_all(sc, vis)
    d = O.shade(os_, cam, 96, 64, vis, T.shade_params(T.SHADE_DEPTH))[m]
    g = np.minimum(F32(1.0), F32(8.0) * depth)
    assert np.array_equal(d[:, 0], g) and np.array_equal(d[:, 1], g) and np.array_equal(d[:, 2], g)
    c = O.shade(os_, cam, 96, 64, vis, T.shade_params(T.SHADE_COLOR))[m]
    lut = sc.objects["lut_idx"][obj].astype(np.int64)
    alb = sc.luts["albedo"][lut * 256 + mat]
    assert np.array_equal(c[:, :3], alb) and np.all(c[:, 3] == 1.0)


def unit_voxel_scene(voxel=(3, 5, 2), translation=(4.0, 4.0, 4.0), rotation=None, albedo=(0.6, 0.3, 0.2),
                     roughness=0.5):
    objs = [{"n": (1, 1, 1), "translation": translation}]
    if rotation is not None:
        objs[0]["rotation"] = rotation
    sc = S.make_scene(objs, bits_per_cluster=[S.voxel_bits([voxel])])
    v = 64 * voxel[2] + 8 * voxel[1] + voxel[0]
    mat = int(sc.materials[v])
    sc.luts["albedo"][mat] = albedo
    sc.luts["roughness"][mat] = roughness
    return sc


@pytest.mark.parametrize("axis,sign", [(0, 1), (0, -1), (1, 1), (1, -1), (2, 1), (2, -1)])
def test_normal_view_on_axis_rays_is_the_entered_face(axis, sign):
    sc = unit_voxel_scene()
    # voxel (3, 5, 2) of a cluster whose min is at the origin (translation 4 - 4)
    centre = np.array([3.5, 5.5, 2.5])
    pos = centre.copy()
    pos[axis] -= sign * 50.0
    cam = S.axis_camera(pos, axis, sign)
    os_ = O.OracleScene(sc)
    vis, _ = O.visibility_culled(os_, cam, 1, 1)
    assert vis[0, 0] != T.VISIBILITY_EMPTY
    out = O.shade(os_, cam, 1, 1, vis, T.shade_params(T.SHADE_NORMAL))[0, 0]
    n = np.zeros(3, np.float32)
    n[axis] = -sign
    assert np.array_equal(out[:3], n * F32(0.5) + F32(0.5))


def face_normal_f64(R, t, voxel_min, o, d):
    """World normal of the face a ray enters a rotated unit voxel through."""
    ol = R.T @ (o - t) + 4.0  # object -> cluster space (n = 1: half extent 4)
    dl = R.T @ d
    dl = dl / np.linalg.norm(dl)
    lo, hi = np.array(voxel_min, float), np.array(voxel_min, float) + 1.0
    with np.errstate(divide="ignore"):
        a, b = (lo - ol) / dl, (hi - ol) / dl
    near = np.minimum(a, b)
    axis = int(np.argmax(near))
    nl = np.zeros(3)
    nl[axis] = -np.sign(dl[axis])
    return R @ nl


def test_normal_view_on_rotated_objects_matches_float64_geometry():
    rng = np.random.default_rng(11)
    for trial in range(30):
        rot = S.random_rotation(rng)
        R = rot.reshape(4, 4).T[:3, :3].astype(np.float64)
        t = rng.uniform(-20, 20, 3)
        vox = tuple(int(v) for v in rng.integers(0, 8, 3))
        sc = unit_voxel_scene(vox, tuple(t), rot)
        centre_cs = np.array(vox, float) + 0.5
        centre = R @ (centre_cs - 4.0) + t
        d = rng.normal(size=3)
        d /= np.linalg.norm(d)
        o = centre - 40.0 * d
        cam = S.direction_camera(o, d)
        os_ = O.OracleScene(sc)
        vis, _ = O.visibility_culled(os_, cam, 1, 1)
        if vis[0, 0] == T.VISIBILITY_EMPTY:
            continue
        out = O.shade(os_, cam, 1, 1, vis, T.shade_params(T.SHADE_NORMAL))[0, 0]
        n = face_normal_f64(R, t, vox, o.astype(np.float32).astype(float), d.astype(np.float32).astype(float))
        np.testing.assert_allclose(out[:3], n * 0.5 + 0.5, atol=1e-5)


def cook_torrance_f64(n, v, l, albedo, metallic, roughness, radiance, ambient):
    """Independent float64 statement of tg_shade plus the ambient term."""
    n, v, l, albedo = (np.asarray(x, float) for x in (n, v, l, albedo))
    h = (v + l) / np.linalg.norm(v + l)
    nh, nl, nv = (float(np.clip(x, 0.0, 1.0)) for x in (h @ n, n @ l, n @ v))
    a2 = roughness ** 4
    D = a2 / (np.pi * (nh * nh * (a2 - 1.0) + 1.0) ** 2)
    F = (1.0 - roughness) + roughness * (1.0 - nv) ** 5
    k = (roughness + 1.0) ** 2 / 8.0
    G = (nl / (nl * (1.0 - k) + k)) * (nv / (nv * (1.0 - k) + k))
    f0 = 0.04 * (1.0 - metallic) + albedo * metallic
    spec = f0 * (D * F * G / max(4.0 * nv * nl, 0.001))
    diff = (1.0 - F) * (1.0 - metallic) * albedo / np.pi
    return ambient * albedo + (diff + spec) * radiance * nl


@pytest.mark.parametrize("view", [T.SHADE_LIT, T.SHADE_SHADING])
def test_lit_views_match_float64_cook_torrance(view):
    rng = np.random.default_rng(5)
    checked = 0
    for trial in range(40):
        albedo = tuple(rng.uniform(0.05, 1.0, 3))
        rough = float(rng.uniform(0.2, 1.0))
        sc = unit_voxel_scene(albedo=albedo, roughness=rough)
        centre = np.array([3.5, 5.5, 2.5])
        if trial < 8:  # axis rays: n.v = 1
            axis, sign = int(rng.integers(0, 3)), float(rng.choice([-1.0, 1.0]))
            pos = centre.copy()
            pos[axis] -= sign * rng.uniform(10.0, 200.0)
            cam = S.axis_camera(pos, axis, sign)
            d = np.eye(3)[axis] * sign
        else:  # oblique rays through the voxel: 0 < n.v < 1
            d = rng.normal(size=3)
            d /= np.linalg.norm(d)
            pos = centre + rng.uniform(-0.3, 0.3, 3) - rng.uniform(10.0, 200.0) * d
            cam = S.direction_camera(pos, d)
        os_ = O.OracleScene(sc)
        vis, _ = O.visibility_culled(os_, cam, 1, 1)
        assert vis[0, 0] != T.VISIBILITY_EMPTY
        light = rng.normal(size=3)
        metallic, radiance, ambient = float(rng.uniform(0, 1)), float(rng.uniform(0.5, 4)), float(rng.uniform(0, .3))
        prm = T.shade_params(view, light_dir=light, light_radiance=radiance, ambient=ambient, metallic=metallic,
                             roughness=-1.0)
        out = O.shade(os_, cam, 1, 1, vis, prm)[0, 0]
        df = np.asarray(d, np.float32).astype(float)
        n = face_normal_f64(np.eye(3), np.array([4.0, 4.0, 4.0]), (3, 5, 2),
                            np.asarray(pos, np.float32).astype(float), df)
        v = -df / np.linalg.norm(df)  # normalize(camera - hit) = -normalize(dir) for depth > 0
        lf = np.asarray(light, np.float32).astype(float)
        lf /= np.linalg.norm(lf)
        alb = np.ones(3) if view == T.SHADE_SHADING else np.asarray(albedo, np.float32).astype(float)
        ref = cook_torrance_f64(n, v, lf, alb, F32(metallic), F32(rough), F32(radiance), F32(ambient))
        np.testing.assert_allclose(out[:3], ref, rtol=2e-5, atol=1e-6)
        assert out[3] == 1.0
        if lf @ n <= 0.0:  # light below the surface: ambient only, exactly
            assert np.array_equal(out[:3], (F32(ambient) * alb.astype(np.float32)).astype(np.float32))
        checked += 1
    assert checked == 40


def test_fixed_roughness_overrides_the_lut_entry():
    sc = unit_voxel_scene(roughness=0.3)
    cam = S.axis_camera((3.5, 5.5, -60.0), 2, 1.0)
    os_ = O.OracleScene(sc)
    vis, _ = O.visibility_culled(os_, cam, 1, 1)
    from_lut = O.shade(os_, cam, 1, 1, vis, T.shade_params(light_dir=(0.2, 0.3, -1.0), roughness=-1.0))
    fixed = O.shade(os_, cam, 1, 1, vis, T.shade_params(light_dir=(0.2, 0.3, -1.0), roughness=0.3))
    other = O.shade(os_, cam, 1, 1, vis, T.shade_params(light_dir=(0.2, 0.3, -1.0), roughness=0.8))
    assert from_lut.tobytes() == fixed.tobytes() and from_lut.tobytes() != other.tobytes()
