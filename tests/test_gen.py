"""Synthetic workloads: determinism, sizes and the statistics the workloads promise."""
from __future__ import annotations

import numpy as np

import tg_amd as T


def test_single_object_workload():
    p = T.single_object_params(1)
    a, b = T.generate_scene(p), T.generate_scene(p)
    assert a.n_clusters == 64 and a.n_cluster_pointers == 64
    assert np.array_equal(a.bits, b.bits) and np.array_equal(a.materials, b.materials)
    assert abs(a.solid_fraction() - 0.3) < 0.02
    o = a.objects[0]
    assert tuple(o["n_cluster_pointers_per_dim"]) == (4, 4, 4)
    assert np.array_equal(o["rotation"], np.eye(4, dtype=np.float32).reshape(-1))
    assert np.all(o["translation"] == 0)
    assert np.array_equal(a.cluster_pointers, np.arange(64))
    lut = a.luts
    assert lut.size == 256 and lut["albedo"].min() >= 0 and lut["albedo"].max() < 1
    c = T.generate_scene(T.single_object_params(2))
    assert not np.array_equal(a.bits, c.bits)


def test_single_object_camera_looks_at_object_centre():
    p = T.single_object_params(1)
    cam = T.generate_camera(p, 0, 256, 256)[0]
    pos = cam["camera"][:3].astype(np.float64)
    assert abs(np.linalg.norm(pos) - 64.0) < 1e-3
    centre_dir = 0.25 * sum(cam[k][:3].astype(np.float64) for k in ("ray_bl", "ray_br", "ray_tr", "ray_tl"))
    centre_dir /= np.linalg.norm(centre_dir)
    assert np.dot(centre_dir, -pos / np.linalg.norm(pos)) > 0.9999
    for k in ("ray_bl", "ray_br", "ray_tr", "ray_tl"):
        assert abs(np.linalg.norm(cam[k][:3]) - 1.0) < 1e-6
    # 60 degree vertical field of view: angle between bottom and top corner
    # rays at the left edge projected on the vertical plane.
    up = cam["ray_tl"][:3] - cam["ray_bl"][:3]
    assert np.linalg.norm(up) > 0


def test_noise_objects_workload():
    p = T.noise_objects_params(seed=1)
    s = T.generate_scene(p)
    assert s.objects.size == 512
    n = s.objects["n_cluster_pointers_per_dim"]
    assert n.min() >= 1 and n.max() <= 8
    assert s.n_clusters == int(np.prod(n, axis=1).sum())
    assert 35_000 < s.n_clusters < 60_000
    assert 0.3 < s.solid_fraction() < 0.5
    assert s.objects["translation"].min() >= 0 and s.objects["translation"].max() < 512
    rot = s.objects["rotation"].reshape(-1, 4, 4).transpose(0, 2, 1)[:, :3, :3].astype(np.float64)
    eye = np.einsum("nij,nkj->nik", rot, rot)
    assert np.abs(eye - np.eye(3)).max() < 1e-5
    assert np.all(np.abs(np.linalg.det(rot) - 1) < 1e-5)
    first = s.objects["first_cluster_pointer"]
    assert first[0] == 0 and np.all(first[1:] == np.cumsum(np.prod(n, axis=1))[:-1])
    assert s.materials.min() < 64 and s.materials.max() > 192


def test_noise_camera_at_world_centre():
    p = T.noise_objects_params(seed=1)
    cam = T.generate_camera(p, 0, 1920, 1080)[0]
    assert np.allclose(cam["camera"][:3], 256.0)
    w = np.linalg.norm(cam["ray_br"][:3] - cam["ray_bl"][:3])
    h = np.linalg.norm(cam["ray_tl"][:3] - cam["ray_bl"][:3])
    assert abs(w / h - 1920 / 1080) < 1e-4
    cams = [T.generate_camera(p, i, 1920, 1080)[0] for i in range(4)]
    assert len({c["ray_bl"].tobytes() for c in cams}) == 4


def test_stream_seeds_are_independent():
    lib = T.load_library()
    seeds = {lib.tgx_gen_stream_seed(1, k) for k in range(1, 9)}
    assert len(seeds) == 8
