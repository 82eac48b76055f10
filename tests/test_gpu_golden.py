"""libtgx.so against the golden fixtures of tests/golden, bit for bit:
config-1 visibility words, the small noise scene's GI frame and SVO
traversal hits (the same fixtures the oracle is held to on the CPU)."""
from __future__ import annotations

import numpy as np
import pytest

import tg_amd as T
from tests import golden_cases as G

pytestmark = pytest.mark.gpu


def test_config1_visibility_matches_fixture(ctx):
    fix = G.load("config1_visibility.npz")
    p, sc, cam, _ = G.make.config1()
    ctx.upload(sc)
    vis = ctx.visibility(G.camera(fix), 256, 256)[0]
    assert np.array_equal(vis, fix["vis"])


def test_noise_gi_and_traces_match_fixtures(ctx):
    fix = G.load("noise_gi.npz")
    rays = G.load("svo_rays.npz")
    p, sc, cam, mn = G.make.noise_case()
    ctx.upload(sc)
    ctx.svo_build(fix["svo_min"], int(fix["depth"]))
    info = ctx.svo_info()
    assert (int(info["n_nodes"]), int(info["n_leaves"])) == (int(fix["svo_nodes"]), int(fix["svo_leaves"]))
    prm = fix["gi_params"].view(T.GI_PARAMS_DTYPE).reshape(1)
    vis, rad, irr = ctx.frame(G.camera(fix), G.make.NOISE_W, G.make.NOISE_H, prm)
    assert np.array_equal(vis, fix["vis"])
    assert np.array_equal(rad.view(np.uint32), fix["radiance_bits"])
    assert np.array_equal(irr.view(np.uint32), fix["irradiance_bits"])
    hits = ctx.svo_trace(rays["origins"], rays["dirs"], float(rays["max_distance"]))
    assert hits.tobytes() == G.hits(rays).tobytes()
