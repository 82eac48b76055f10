"""GPU parity of the SVO: build (canonical export) and traversal against
the CPU oracle, bit for bit, through the C ABI."""
from __future__ import annotations

import sys

import numpy as np
import pytest

import oracle as O
import tg_amd as T
from tests import scenes as S

pytestmark = pytest.mark.gpu


def say(msg: str) -> None:
    print(f"[gpu] {msg}", file=sys.stderr, flush=True)


def assert_same_svo(gpu: dict, ref: O.SvoExport, what: str) -> None:
    gi, ri = gpu["info"], ref.info
    for f in ("depth", "n_nodes", "n_leaves", "n_list_entries", "flags"):
        assert gi[f] == ri[f], f"{what}: info.{f} {gi[f]} != {ri[f]}"
    assert np.array_equal(gpu["nodes"], ref.nodes), f"{what}: nodes differ"
    bad = np.nonzero((gpu["bits"] != ref.bits).any(axis=1))[0]
    assert bad.size == 0, f"{what}: leaf bits differ in {bad.size} leaves, first {bad[:5]}"
    bad = np.nonzero((gpu["materials"] != ref.materials).any(axis=1))[0]
    assert bad.size == 0, f"{what}: leaf materials differ in {bad.size} leaves"
    assert np.array_equal(gpu["leaf_n"], ref.leaf_n), f"{what}: leaf n differ"
    assert np.array_equal(gpu["list_offsets"], ref.list_offsets), f"{what}: list offsets differ"
    n = int(ri["n_list_entries"])
    assert np.array_equal(gpu["list"][:n], ref.list[:n]), f"{what}: lists differ"


def assert_same_hits(gpu: np.ndarray, ref: np.ndarray, what: str) -> None:
    for f in ("hit", "block", "voxel", "normal", "steps"):
        bad = np.nonzero((gpu[f] != ref[f]).reshape(len(gpu), -1).any(axis=1))[0]
        assert bad.size == 0, f"{what}: field {f} differs for {bad.size} rays, first {bad[:5]}"
    d = gpu["distance"].view(np.uint32) != ref["distance"].view(np.uint32)
    assert not d.any(), f"{what}: distance bits differ for {int(d.sum())} rays"


def rays_from(rng, n, lo, hi, centre=None):
    o = rng.uniform(lo, hi, size=(n, 3)).astype(np.float32)
    if centre is not None:
        o[: n // 2] = np.asarray(centre, np.float32)
    d = rng.normal(size=(n, 3))
    d = (d / np.linalg.norm(d, axis=1, keepdims=True)).astype(np.float32)
    d[:8] = np.eye(3, dtype=np.float32)[np.arange(8) % 3] * np.where(np.arange(8) % 2, 1, -1)[:, None]
    return o, d


@pytest.mark.parametrize("flags", [0, T.SVO_FLAG_REFERENCE_CORNER_QUIRK])
def test_svo_small_scenes_bit_exact(ctx, flags):
    rng = np.random.default_rng(7)
    objs = [{"n": (2, 1, 3), "translation": (20.0, 40.0, 36.0)},
            {"n": (3, 3, 1), "translation": (76.0, 60.0, 100.0)}]
    objs += [{"n": tuple(rng.integers(1, 4, 3)), "translation": tuple(rng.uniform(10, 118, 3)),
              "rotation": S.random_rotation(rng)} for _ in range(12)]
    sc = S.make_scene(objs, seed=3)
    ctx.upload(sc)
    for depth, mn in ((7, (0.0, 0.0, 0.0)), (8, (-64.0, -32.0, -96.0))):
        say(f"small scene SVO depth {depth} flags {flags}")
        ctx.svo_build(mn, depth, flags)
        ref = O.OracleSvo(O.OracleScene(sc), mn, depth, flags)
        assert_same_svo(ctx.svo_export(), ref.export(), f"depth {depth}")
        o, d = rays_from(rng, 4000, -10, 140, centre=(64.0, 64.0, 64.0))
        say(f"small scene trace depth {depth}")
        assert_same_hits(ctx.svo_trace(o, d, 300.0), ref.trace(o, d, 300.0), f"depth {depth}")


def test_svo_noise_scene_depth11_bit_exact(ctx):
    p = T.noise_objects_params(seed=1)
    sc = T.generate_scene(p)
    ctx.upload(sc)
    cam = T.generate_camera(p, 0, 1920, 1080)[0]
    mn = np.floor(cam["camera"][:3]) - 1024.0
    say("512-object scene SVO depth 11")
    ctx.svo_build(mn, 11)
    ref = O.OracleSvo(O.OracleScene(sc), mn, 11)
    assert_same_svo(ctx.svo_export(), ref.export(), "depth 11")
    rng = np.random.default_rng(1)
    o, d = rays_from(rng, 20000, 0, 512, centre=cam["camera"][:3])
    say("512-object scene trace 20000 rays")
    g = ctx.svo_trace(o, d, 1024.0)
    assert_same_hits(g, ref.trace(o, d, 1024.0), "depth 11 trace")
    assert 0.2 < g["hit"].mean() < 0.95


def test_svo_error_paths(ctx):
    sc = T.generate_scene(T.single_object_params(1))
    ctx.upload(sc)
    for mn, depth in (((0.5, 0.0, 0.0), 8), ((0.0, 0.0, 0.0), 5), ((0.0, 0.0, 0.0), 14)):
        with pytest.raises(T.TgxError) as e:
            ctx.svo_build(mn, depth)
        assert e.value.code == 1
    fresh = T.Context(0)
    try:
        with pytest.raises(T.TgxError) as e:
            fresh.svo_trace(np.zeros((1, 3)), np.ones((1, 3)))
        assert e.value.code == 7
    finally:
        fresh.close()


def svo_visibility_reference(svo: O.OracleSvo, cam: np.ndarray, w: int, h: int) -> np.ndarray:
    """debug_visibility_svo.frag:27-70 over the oracle traversal: rays in the
    kernels' fp32 operation order, words packed as the reference packs them."""
    f = np.float32
    c = cam[0]
    ys, xs = np.mgrid[0:h, 0:w]
    fx = (xs.astype(f) + f(0.5)) / f(w)
    fy = f(1.0) - (ys.astype(f) + f(0.5)) / f(h)
    omy, omx = f(1.0) - fy, f(1.0) - fx
    d = np.empty((h, w, 3), f)
    for a in range(3):
        left = c["ray_bl"][a] * omy + c["ray_tl"][a] * fy
        right = c["ray_br"][a] * omy + c["ray_tr"][a] * fy
        d[..., a] = left * omx + right * fx
    m = np.sqrt(d[..., 0] * d[..., 0] + d[..., 1] * d[..., 1] + d[..., 2] * d[..., 2])
    d = d / m[..., None]
    o = np.broadcast_to(c["camera"][:3], d.shape)
    hits = svo.trace(o.reshape(-1, 3), d.reshape(-1, 3))
    e = svo.export()
    leaf_level = int(svo.info()["depth"]) - 5
    index = {(int(n["level"]),) + tuple(int(v) for v in n["coord"]): i for i, n in enumerate(e.nodes)}
    out = np.full(w * h, T.VISIBILITY_EMPTY, np.uint64)
    far = f(c["far_plane"])
    for i in np.nonzero(hits["hit"])[0]:
        dd = f(hits["distance"][i]) / far
        if dd <= f(1.0):
            node = index[(leaf_level,) + tuple(int(v) for v in hits["block"][i])]
            out[i] = (int(f(dd * f(16777215.0))) << 40) | ((node & 0x7FFFFFFF) << 9) | (int(hits["voxel"][i]) % 512)
    return out.reshape(h, w)


@pytest.mark.parametrize("w,h", [(160, 90), (480, 270)])
def test_svo_visibility_matches_oracle(ctx, w, h):
    p = T.noise_objects_params(seed=1)
    sc = T.generate_scene(p)
    cam = T.generate_camera(p, 0, w, h)
    mn = np.floor(cam[0]["camera"][:3]) - 1024.0
    ctx.upload(sc)
    ctx.svo_build(mn, 11)
    say(f"svo visibility {w}x{h}")
    got = ctx.svo_visibility(cam, w, h)
    svo = O.OracleSvo(O.OracleScene(sc), mn, 11)
    ref = svo_visibility_reference(svo, cam, w, h)
    bad = got != ref
    assert not bad.any(), f"{int(bad.sum())} words differ, first at {np.argwhere(bad)[:3].tolist()}"
    assert (got != T.VISIBILITY_EMPTY).mean() > 0.3
