"""Writes the golden fixtures of this directory from the oracle.

    python tests/golden/make_golden.py

The reference holds no vectors, fixtures or known-answer tests for its
visibility/SVO path, so these fixtures are the oracle's outputs frozen at
the time they were written. Each file stores its inputs (generator
parameters, camera records, rays, GI parameters), a SHA-256 of the
generated scene arrays (so a change of the scene generator is reported as
such) and the outputs:

- config1_visibility.npz: the single-object scene (4x4x4 clusters,
  Bernoulli(0.3) solids, seed 1), 256x256 camera 0: every 64-bit
  visibility word (visibility.frag:194-206 packing).
- noise_gi.npz: 30 noise objects of 1-5 clusters per axis in a 96^3 world
  (seed 6), 64x48 camera 0, SVO of depth 7 around the camera, one-bounce GI
  at 4 spp: visibility words, SVO counts and a SHA-256 of its canonical
  export, radiance and irradiance as float32 bit patterns, ray counts.
- svo_rays.npz: 512 seeded rays through that SVO (origins in the world box,
  unit directions, maximum distance 300) and the traversal's hit records
  (hit flag, distance bit pattern, leaf block, voxel, face normal, node
  steps) as raw bytes.
"""
from __future__ import annotations

import hashlib
import sys
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
ROOT = HERE.parent.parent
if str(ROOT) not in sys.path:
    sys.path.insert(0, str(ROOT))

import oracle as O  # noqa: E402
import tg_amd as T  # noqa: E402


def scene_sha256(sc) -> str:
    h = hashlib.sha256()
    for a in (sc.objects, sc.cluster_pointers, sc.cluster_idx_to_object, sc.bits, sc.materials, sc.luts):
        h.update(np.ascontiguousarray(a).tobytes())
    return h.hexdigest()


def export_sha256(e) -> str:
    h = hashlib.sha256()
    for a in (e.nodes, e.bits, e.materials, e.leaf_n, e.list_offsets, e.list):
        h.update(np.ascontiguousarray(a).tobytes())
    return h.hexdigest()


def config1():
    p = T.single_object_params(1)
    sc = T.generate_scene(p)
    cam = T.generate_camera(p, 0, 256, 256)
    vis = O.visibility_brute(O.OracleScene(sc), cam, 256, 256)
    return p, sc, cam, vis


NOISE = dict(seed=6, n_objects=30, max_extent=5, world=96.0)
NOISE_W, NOISE_H, NOISE_DEPTH = 64, 48, 7


def noise_case():
    p = T.noise_objects_params(**NOISE)
    sc = T.generate_scene(p)
    cam = T.generate_camera(p, 0, NOISE_W, NOISE_H)
    mn = (np.floor(cam[0]["camera"][:3]) - float(1 << (NOISE_DEPTH - 1))).astype(np.float32)
    return p, sc, cam, mn


def noise_gi_params():
    return T.gi_params(T.seeded_sun(6), spp=4, frame=3)


RAY_MAX_DISTANCE = np.float32(300.0)


def golden_rays(mn: np.ndarray, n: int = 512):
    rng = np.random.default_rng(20261203)
    side = float(1 << NOISE_DEPTH)
    o = (mn + rng.uniform(0.0, side, (n, 3))).astype(np.float32)
    d = rng.normal(size=(n, 3))
    d = (d / np.linalg.norm(d, axis=1, keepdims=True)).astype(np.float32)
    return o, d


def main() -> None:
    p, sc, cam, vis = config1()
    np.savez_compressed(HERE / "config1_visibility.npz", scene_sha256=np.array(scene_sha256(sc)),
                        camera=cam.view(np.uint8), width=256, height=256, vis=vis)

    p, sc, cam, mn = noise_case()
    os_ = O.OracleScene(sc)
    vis, _ = O.visibility_culled(os_, cam, NOISE_W, NOISE_H)
    svo = O.OracleSvo(os_, mn, NOISE_DEPTH)
    info = svo.info()
    prm = noise_gi_params()
    rad, irr, cnt = svo.gi(cam, NOISE_W, NOISE_H, vis, prm)
    np.savez_compressed(HERE / "noise_gi.npz", scene_sha256=np.array(scene_sha256(sc)), camera=cam.view(np.uint8),
                        svo_min=mn, depth=NOISE_DEPTH, gi_params=prm.view(np.uint8), vis=vis,
                        svo_nodes=int(info["n_nodes"]), svo_leaves=int(info["n_leaves"]),
                        svo_export_sha256=np.array(export_sha256(svo.export())),
                        radiance_bits=rad.view(np.uint32), irradiance_bits=irr.view(np.uint32), counts=cnt)

    o, d = golden_rays(mn)
    hits = svo.trace(o, d, RAY_MAX_DISTANCE)
    np.savez_compressed(HERE / "svo_rays.npz", scene_sha256=np.array(scene_sha256(sc)), origins=o, dirs=d,
                        max_distance=RAY_MAX_DISTANCE, hits=hits.view(np.uint8).reshape(len(hits), -1))
    for f in sorted(HERE.glob("*.npz")):
        print(f"{f.name}: {f.stat().st_size} bytes")


if __name__ == "__main__":
    main()
