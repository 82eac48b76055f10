"""The oracle against the golden fixtures of tests/golden (written by
tests/golden/make_golden.py): the scenes the generator makes today are the
fixtures' scenes, and the oracle's visibility words, SVO, GI radiance and
traversal hits are the stored ones bit for bit. A change to the word
packing (visibility.frag:194-206), the traversal or the estimator that
the oracle and the kernels make together fails here."""
from __future__ import annotations

import numpy as np

import oracle as O
import tg_amd as T
from tests import golden_cases as G


def test_config1_visibility_words():
    fix = G.load("config1_visibility.npz")
    p, sc, cam, vis = G.make.config1()
    assert G.make.scene_sha256(sc) == str(fix["scene_sha256"])
    assert cam.tobytes() == G.camera(fix).tobytes()
    assert np.array_equal(vis, fix["vis"])
    # the stored words decode to hits in front of the camera at ~64 voxels
    hits = fix["vis"][fix["vis"] != T.VISIBILITY_EMPTY]
    assert 0.2 < hits.size / fix["vis"].size < 0.5
    depth = (hits >> np.uint64(40)).astype(np.float64) / 16777215.0 * float(cam[0]["far_plane"])
    assert depth.min() > 40.0 and depth.max() < 90.0
    assert int(((hits >> np.uint64(9)) & np.uint64(0x7FFFFFFF)).max()) < sc.n_cluster_pointers


def test_noise_scene_svo_and_gi():
    fix = G.load("noise_gi.npz")
    p, sc, cam, mn = G.make.noise_case()
    assert G.make.scene_sha256(sc) == str(fix["scene_sha256"])
    os_ = O.OracleScene(sc)
    vis, _ = O.visibility_culled(os_, cam, G.make.NOISE_W, G.make.NOISE_H)
    assert np.array_equal(vis, fix["vis"])
    svo = O.OracleSvo(os_, mn, G.make.NOISE_DEPTH)
    info = svo.info()
    assert (int(info["n_nodes"]), int(info["n_leaves"])) == (int(fix["svo_nodes"]), int(fix["svo_leaves"]))
    assert G.make.export_sha256(svo.export()) == str(fix["svo_export_sha256"])
    prm = G.make.noise_gi_params()
    assert prm.tobytes() == fix["gi_params"].tobytes()
    rad, irr, cnt = svo.gi(cam, G.make.NOISE_W, G.make.NOISE_H, vis, prm)
    assert np.array_equal(rad.view(np.uint32), fix["radiance_bits"])
    assert np.array_equal(irr.view(np.uint32), fix["irradiance_bits"])
    assert np.array_equal(cnt, fix["counts"])
    assert (fix["radiance_bits"] != 0).any()


def test_svo_ray_hits():
    fix = G.load("svo_rays.npz")
    p, sc, cam, mn = G.make.noise_case()
    assert G.make.scene_sha256(sc) == str(fix["scene_sha256"])
    o, d = G.make.golden_rays(mn)
    assert np.array_equal(o, fix["origins"]) and np.array_equal(d, fix["dirs"])
    svo = O.OracleSvo(O.OracleScene(sc), mn, G.make.NOISE_DEPTH)
    got = svo.trace(o, d, float(fix["max_distance"]))
    assert got.tobytes() == G.hits(fix).tobytes()
    assert 0.1 < G.hits(fix)["hit"].mean() < 0.9
