"""GPU shading pass (tgx_shade_device) against the oracle, bit for bit, in
every view: the decode, hash colours, normals and GGX lighting use the same
fp32 operations in the same order on both sides."""
from __future__ import annotations

import sys

import numpy as np
import pytest

import oracle as O
import tg_amd as T

pytestmark = pytest.mark.gpu

VIEWS = list(range(10))


def say(msg: str) -> None:
    print(f"[gpu] {msg}", file=sys.stderr, flush=True)


@pytest.mark.parametrize("scene,w,h", [("single", 256, 256), ("noise", 480, 270)])
def test_shade_views_bit_exact(ctx, scene, w, h):
    p = T.single_object_params(1) if scene == "single" else T.noise_objects_params(seed=1, n_luts=4)
    sc = T.generate_scene(p)
    cam = T.generate_camera(p, 0, w, h)
    ctx.upload(sc)
    vis = ctx.visibility(cam, w, h)[0]
    os_ = O.OracleScene(sc)
    assert (vis != T.VISIBILITY_EMPTY).mean() > 0.05
    for view in VIEWS:
        say(f"shade {scene} {w}x{h} view {view}")
        prm = T.shade_params(view, light_dir=(0.3, 0.8, -0.4), roughness=-1.0 if view == 0 else 0.8)
        got = ctx.shade(cam, w, h, vis, prm)
        ref = O.shade(os_, cam, w, h, vis, prm)
        assert got.tobytes() == ref.tobytes(), f"view {view}: {int((got != ref).any(-1).sum())} pixels differ"


def test_shade_rejects_bad_arguments(ctx):
    p = T.single_object_params(1)
    ctx.upload(T.generate_scene(p))
    cam = T.generate_camera(p, 0, 8, 8)
    vis = np.full((8, 8), T.VISIBILITY_EMPTY, np.uint64)
    with pytest.raises(T.TgxError):
        ctx.shade(cam, 8, 8, vis, T.shade_params(10))
    with pytest.raises(T.TgxError):
        ctx.shade(cam, 8, 8, vis, T.shade_params(light_dir=(0.0, 0.0, 0.0)))
    out = ctx.shade(cam, 8, 8, vis, T.shade_params())
    assert np.all(out == np.array([1.0, 0.0, 1.0, 1.0], np.float32))
