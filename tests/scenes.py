"""Hand-built scenes and cameras for closed-form tests."""
from __future__ import annotations

import numpy as np

import tg_amd as T


def identity_rotation() -> np.ndarray:
    r = np.zeros(16, np.float32)
    r[[0, 5, 10, 15]] = 1.0
    return r


def rotation_from_matrix(m3: np.ndarray) -> np.ndarray:
    r = np.zeros(16, np.float32)
    for row in range(3):
        for col in range(3):
            r[4 * col + row] = np.float32(m3[row, col])
    r[15] = 1.0
    return r


def random_rotation(rng: np.random.Generator) -> np.ndarray:
    q = rng.normal(size=4)
    q /= np.linalg.norm(q)
    x, y, z, w = q
    m = np.array([
        [1 - 2 * (y * y + z * z), 2 * (x * y - z * w), 2 * (x * z + y * w)],
        [2 * (x * y + z * w), 1 - 2 * (x * x + z * z), 2 * (y * z - x * w)],
        [2 * (x * z - y * w), 2 * (y * z + x * w), 1 - 2 * (x * x + y * y)],
    ])
    return rotation_from_matrix(m)


def make_scene(objects: list[dict], bits_per_cluster: list[np.ndarray] | None = None,
               cluster_perm: np.ndarray | None = None, seed: int = 0) -> T.Scene:
    """objects: dicts with n (3 ints), translation, rotation (16 floats).
    bits_per_cluster: one u32[16] per cluster pointer (default random 30%).
    cluster_perm: cluster idx of each pointer (default identity)."""
    rng = np.random.default_rng(seed)
    objs = np.zeros(len(objects), T.OBJECT_DTYPE)
    n_cptr = 0
    owner = []
    for i, o in enumerate(objects):
        n = np.asarray(o["n"], np.uint32)
        objs[i]["n_cluster_pointers_per_dim"] = n
        objs[i]["first_cluster_pointer"] = n_cptr
        objs[i]["translation"] = np.asarray(o.get("translation", (0, 0, 0)), np.float32)
        objs[i]["rotation"] = np.asarray(o.get("rotation", identity_rotation()), np.float32)
        objs[i]["lut_idx"] = 0
        cnt = int(n.prod())
        owner += [i] * cnt
        n_cptr += cnt
    perm = np.arange(n_cptr, dtype=np.uint32) if cluster_perm is None else np.asarray(cluster_perm, np.uint32)
    bits = np.zeros(16 * n_cptr, np.uint32)
    c2o = np.zeros(n_cptr, np.uint32)
    mats = np.zeros(512 * n_cptr, np.uint8)
    for cptr in range(n_cptr):
        cidx = int(perm[cptr])
        if bits_per_cluster is None:
            b = np.packbits(rng.random(512) < 0.3, bitorder="little").view(np.uint32)
        else:
            b = np.asarray(bits_per_cluster[cptr], np.uint32)
        bits[16 * cidx:16 * cidx + 16] = b
        c2o[cidx] = owner[cptr]
        mats[512 * cidx:512 * cidx + 512] = rng.integers(0, 256, 512, dtype=np.uint8)
    luts = np.zeros(256, T.LUT_DTYPE)
    luts["albedo"] = rng.random((256, 3), dtype=np.float32)
    luts["roughness"] = rng.random(256, dtype=np.float32)
    return T.Scene(objs, perm.copy(), c2o, bits, mats, luts)


def voxel_bits(voxels: list[tuple[int, int, int]]) -> np.ndarray:
    """u32[16] with the given (x, y, z) voxels of one cluster set."""
    b = np.zeros(16, np.uint32)
    for x, y, z in voxels:
        v = 64 * z + 8 * y + x
        b[v >> 5] |= np.uint32(1) << np.uint32(v & 31)
    return b


def axis_camera(position, forward_axis: int, sign: float, far: float = 1000.0) -> np.ndarray:
    """Camera whose 1x1-pixel ray is exactly sign * e_axis: symmetric corner
    rays around the axis, so the pixel-centre lerp (fx = fy = 0.5) cancels
    the lateral parts exactly."""
    a = [i for i in range(3) if i != forward_axis]
    cam = np.zeros(1, T.CAMERA_DTYPE)
    cam["camera"][0, :3] = position
    corners = {"ray_bl": (-1, -1), "ray_br": (1, -1), "ray_tr": (1, 1), "ray_tl": (-1, 1)}
    for name, (u, v) in corners.items():
        d = np.zeros(4, np.float32)
        d[forward_axis] = np.float32(sign * 0.75)
        d[a[0]] = np.float32(0.5 * u)
        d[a[1]] = np.float32(0.5 * v)
        cam[name][0] = d
    cam["near_plane"] = 0.1
    cam["far_plane"] = far
    return cam


def direction_camera(position, direction, far: float = 1000.0) -> np.ndarray:
    """Camera whose four corner rays all equal `direction`: the 1x1 pixel
    (fx = fy = 0.5) then has exactly that direction (mix of equal values at
    t = 0.5 is exact)."""
    cam = np.zeros(1, T.CAMERA_DTYPE)
    cam["camera"][0, :3] = position
    for name in ("ray_bl", "ray_br", "ray_tr", "ray_tl"):
        cam[name][0, :3] = np.asarray(direction, np.float32)
    cam["near_plane"] = 0.1
    cam["far_plane"] = far
    return cam


def permute_clusters(scene: T.Scene, seed: int) -> T.Scene:
    """Same scene with cluster indices permuted (bits, materials and the
    cluster->object map move with them)."""
    rng = np.random.default_rng(seed)
    n = scene.n_clusters
    new_of_old = rng.permutation(n).astype(np.uint32)
    bits = np.zeros_like(scene.bits)
    mats = np.zeros_like(scene.materials)
    c2o = np.zeros_like(scene.cluster_idx_to_object)
    b16 = scene.bits.reshape(n, 16)
    m512 = scene.materials.reshape(n, 512)
    bits.reshape(n, 16)[new_of_old] = b16
    mats.reshape(n, 512)[new_of_old] = m512
    c2o[new_of_old] = scene.cluster_idx_to_object
    cptrs = new_of_old[scene.cluster_pointers]
    return T.Scene(scene.objects.copy(), cptrs, c2o, bits, mats, scene.luts.copy())
