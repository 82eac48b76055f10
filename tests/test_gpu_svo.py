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


def fnv1a64(data: bytes) -> int:
    h = 0xCBF29CE484222325
    for b in data:
        h = ((h ^ b) * 0x100000001B3) & 0xFFFFFFFFFFFFFFFF
    return h


def resealed(data: bytearray) -> bytes:
    """The file with its trailing FNV-1a 64 checksum recomputed."""
    body = bytes(data[:-8])
    return body + np.uint64(fnv1a64(body)).tobytes()


SVO_HEADER_BYTES = 64   # magic, version, n_inner, tgx_svo_info (32 B), n_clusters, n_luts, n_solid
SVO_NODE_BYTES = 32


def test_svo_file_round_trip(ctx, tmp_path):
    """tgx_svo_save / tgx_svo_load: a loaded SVO exports, traces and renders
    GI exactly as the built one, and equals the oracle's build."""
    p = T.single_object_params(1)
    sc = T.generate_scene(p)
    cam = T.generate_camera(p, 0, 96, 64)
    mn = np.array([-64.0, -64.0, -64.0], np.float32)
    prm = T.gi_params(T.seeded_sun(1), spp=2)
    ctx.upload(sc)
    say("SVO file: build depth 7, save")
    ctx.svo_build(mn, 7)
    built = ctx.svo_export()
    f1, f2 = tmp_path / "a.tgxsvo", tmp_path / "b.tgxsvo"
    ctx.svo_save(f1)
    vis, rad, irr = ctx.frame(cam, 96, 64, prm)
    rng = np.random.default_rng(5)
    o, d = rays_from(rng, 3000, -80, 80, centre=(0.0, 0.0, 0.0))
    hits = ctx.svo_trace(o, d, 300.0)
    with T.Context(0) as c2:
        c2.upload(sc)
        say("SVO file: load into a fresh context")
        c2.svo_load(f1)
        loaded = c2.svo_export()
        ref = O.OracleSvo(O.OracleScene(sc), mn, 7).export()
        assert_same_svo(loaded, ref, "loaded vs oracle")
        for k in ("nodes", "bits", "materials", "leaf_n", "list_offsets", "list"):
            assert np.array_equal(loaded[k], built[k]), f"loaded {k} differs from the build"
        assert loaded["info"].tobytes() == built["info"].tobytes()
        assert_same_hits(c2.svo_trace(o, d, 300.0), hits, "loaded trace")
        v2, r2, i2 = c2.frame(cam, 96, 64, prm)
        assert v2.tobytes() == vis.tobytes() and r2.tobytes() == rad.tobytes() and i2.tobytes() == irr.tobytes()
        c2.svo_save(f2)
        assert f1.read_bytes() == f2.read_bytes(), "save of a loaded SVO is not the same file"
        say("SVO file: update after load samples every leaf")
        reused, vox = c2.svo_update()
        assert reused == 0 and vox == int(built["info"]["n_leaves"])
        assert_same_svo(c2.svo_export(), ref, "update after load")
    assert rad.any()


def test_svo_file_rejects_bad_files(ctx, tmp_path):
    """Corrupt, truncated, foreign and mismatched files fail with an error
    before anything reaches the device, and the context keeps its SVO."""
    objs = [{"n": (2, 2, 2), "translation": (40.0, 40.0, 40.0)},
            {"n": (1, 3, 2), "translation": (90.0, 70.0, 60.0)}]
    sc = S.make_scene(objs, seed=4)
    ctx.upload(sc)
    ctx.svo_build((0.0, 0.0, 0.0), 8)
    good = ctx.svo_export()
    path = tmp_path / "s.tgxsvo"
    ctx.svo_save(path)
    raw = bytearray(path.read_bytes())
    info = good["info"]
    n_nodes, nl = int(info["n_nodes"]), int(info["n_leaves"])
    n_inner = n_nodes - nl
    assert nl > 0 and int(info["n_list_entries"]) > 0

    def expect(data: bytes, code: int, words: str, name: str) -> None:
        q = tmp_path / name
        q.write_bytes(data)
        with pytest.raises(T.TgxError) as e:
            ctx.svo_load(q)
        assert e.value.code == code, (name, str(e.value))
        assert words in str(e.value), (name, str(e.value))

    say("SVO file: rejected files")
    flipped = bytearray(raw)
    flipped[len(raw) // 2] ^= 0x10
    expect(bytes(flipped), 1, "", "flipped")
    expect(bytes(raw[:-100]), 1, "", "truncated")
    expect(bytes(raw) + b"\0", 1, "", "trailing")
    expect(b"TGXSCN01" + bytes(raw[8:]), 1, "not an SVO file", "magic")
    # child range past the node array, with a valid checksum
    bad = bytearray(raw)
    bad[SVO_HEADER_BYTES:SVO_HEADER_BYTES + 4] = np.uint32(n_nodes + 5).tobytes()
    expect(resealed(bad), 1, "bad SVO nodes", "child_range")
    # a leaf-level child that points back into the inner nodes
    bad = bytearray(raw)
    last_inner = SVO_HEADER_BYTES + SVO_NODE_BYTES * (n_inner - 1)
    bad[last_inner:last_inner + 4] = np.uint32(1).tobytes()
    expect(resealed(bad), 1, "bad SVO nodes", "leaf_child")
    # a cluster index past the scene's clusters
    lst = SVO_HEADER_BYTES + SVO_NODE_BYTES * n_nodes + 4 * nl + 4 * (nl + 1)
    bad = bytearray(raw)
    bad[lst:lst + 4] = np.uint32(0x7FFFFFFF).tobytes()
    expect(resealed(bad), 1, "cluster index", "cluster")
    # the root's level
    bad = bytearray(raw)
    bad[SVO_HEADER_BYTES + 12:SVO_HEADER_BYTES + 16] = np.uint32(int.from_bytes(
        raw[SVO_HEADER_BYTES + 12:SVO_HEADER_BYTES + 16], "little") | (1 << 16)).tobytes()
    expect(resealed(bad), 1, "bad SVO nodes", "root_level")
    # a material whose LUT index is past the scene's LUTs (leaf data: no SVO left)
    first_mat = lst + 4 * int(info["n_list_entries"])
    for leaf_bits in good["bits"]:
        first_mat += 4096
        n_solid = int(np.unpackbits(leaf_bits.view(np.uint8)).sum())
        if n_solid:
            break
        first_mat += 2 * n_solid
    assert n_solid
    bad = bytearray(raw)
    bad[first_mat:first_mat + 2] = np.uint16(0xFF00).tobytes()
    assert sc.luts.size // 256 < 255
    expect(resealed(bad), 1, "bad leaf data", "material")
    # every rejection above came before the device was touched
    assert ctx.svo_info().tobytes() == info.tobytes()
    for k in ("nodes", "bits", "materials", "list"):
        assert np.array_equal(ctx.svo_export()[k], good[k])
    with pytest.raises(T.TgxError) as e:
        ctx.svo_load(tmp_path / "missing.tgxsvo")
    assert e.value.code == 8
    say("SVO file: scene mismatch")
    ctx.upload(S.make_scene(objs[:1], seed=4))
    expect(bytes(raw), 1, "clusters", "other_scene")
    ctx.upload(sc)
    ctx.svo_load(path)
    for k in ("nodes", "bits", "materials", "list"):
        assert np.array_equal(ctx.svo_export()[k], good[k])


def test_svo_depth13_far_corner_bit_exact(ctx):
    """Depth 13 (root side 8192): inner nodes sit at levels up to 7, whose
    coordinates reach 127, the largest value the traversal's 7-bit node
    fields hold. Objects at the root's far corner, its centre and its min
    corner; rays around each."""
    rng = np.random.default_rng(13)
    objs = []
    for base in (8150.0, 4090.0, 40.0):
        objs += [{"n": tuple(rng.integers(1, 4, 3)), "translation": tuple(base + rng.uniform(-24, 24, 3)),
                  "rotation": S.random_rotation(rng)} for _ in range(5)]
    objs.append({"n": (2, 2, 2), "translation": (8180.0, 8180.0, 8180.0)})
    sc = S.make_scene(objs, seed=13)
    ctx.upload(sc)
    mn = (0.0, 0.0, 0.0)
    say("depth-13 SVO, objects at the far corner")
    ctx.svo_build(mn, 13)
    ref = O.OracleSvo(O.OracleScene(sc), mn, 13)
    e = ref.export()
    assert_same_svo(ctx.svo_export(), e, "depth 13")
    inner = e.nodes[e.nodes["level"] == 7]
    assert inner.size and int(inner["coord"].max()) == 127
    for base in (8150.0, 4090.0, 40.0):
        o, d = rays_from(rng, 3000, base - 60, base + 60, centre=(base, base, base))
        o = np.clip(o, 0.5, 8191.5).astype(np.float32)
        say(f"depth-13 trace around {base}")
        assert_same_hits(ctx.svo_trace(o, d, 200.0), ref.trace(o, d, 200.0), f"depth 13 at {base}")


def test_svo_depth12_wide_world_bit_exact(ctx):
    """Depth 12 over the batched workload's world shape: objects of 1-32
    clusters per axis in a 4096^3 world (24 of them, so the oracle builds it
    in seconds), the root centred on the first camera. SVO, 20000 traced
    rays and a GI frame bit for bit."""
    p = T.noise_objects_params(seed=12, n_objects=24, max_extent=32, world=4096.0, n_luts=64)
    ctx.generate_scene_device(p)  # equals the host generator's scene (test_gpu_frame)
    sc = ctx.scene_export(p)
    cam = T.generate_camera(p, 0, 480, 270)
    mn = np.floor(cam[0]["camera"][:3]) - 2048.0
    say("24-object 4096^3 world, SVO depth 12")
    ctx.svo_build(mn, 12)
    osc = O.OracleScene(sc)
    ref = O.OracleSvo(osc, mn, 12)
    assert_same_svo(ctx.svo_export(), ref.export(), "depth 12")
    rng = np.random.default_rng(12)
    centres = sc.objects["translation"][:8]
    o = np.concatenate([c + rng.uniform(-150, 150, (2500, 3)) for c in centres]).astype(np.float32)
    d = rng.normal(size=o.shape)
    d = (d / np.linalg.norm(d, axis=1, keepdims=True)).astype(np.float32)
    say("depth-12 trace 20000 rays")
    g = ctx.svo_trace(o, d, 1024.0)
    assert_same_hits(g, ref.trace(o, d, 1024.0), "depth 12 trace")
    assert g["hit"].mean() > 0.05
    prm = T.gi_params(T.seeded_sun(12), spp=2)
    say("depth-12 GI frame 480x270")
    vis, rad, irr = ctx.frame(cam, 480, 270, prm)
    rr, ri, _ = ref.gi(cam, 480, 270, vis, prm)
    assert (vis != T.VISIBILITY_EMPTY).mean() > 0.02
    assert rad.tobytes() == rr.tobytes() and irr.tobytes() == ri.tobytes()


def test_svo_trace_dda_forms_bit_exact(ctx):
    """The leaf DDA's forms (counted; uncounted; uncounted with exact
    crossings of 8^3 bricks without solid voxels at several wave thresholds),
    1 to 4 node-loop iterations per lane step, and queue slices of 0 to 256
    rays per cursor atomic with refill thresholds of 1 to 64 idle lanes, on
    one or two workgroups per CU (trace_small), give the oracle's hits on
    rays chosen to stress the crossing: unit
    directions, directions with equal-magnitude components (ties between
    axes), components of 1e-6, 1e-30 and 0, and directions scaled by 1e-3
    to 1e3."""
    p = T.noise_objects_params(seed=1)
    sc = T.generate_scene(p)
    ctx.upload(sc)
    cam = T.generate_camera(p, 0, 1920, 1080)[0]
    mn = np.floor(cam["camera"][:3]) - 1024.0
    ctx.svo_build(mn, 11)
    ref_svo = O.OracleSvo(O.OracleScene(sc), mn, 11)
    rng = np.random.default_rng(5)
    o, d = rays_from(rng, 24000, 0, 512, centre=cam["camera"][:3])
    d = d.astype(np.float64)
    k = np.arange(len(d))
    sgn = np.where(rng.uniform(size=d.shape) < 0.5, -1.0, 1.0)
    diag = k % 6 == 1
    d[diag] = sgn[diag] * np.abs(d[diag, :1])                      # |dx| = |dy| = |dz|
    pair = k % 6 == 2
    d[pair, 1] = sgn[pair, 1] * np.abs(d[pair, 0])                 # |dx| = |dy|
    tiny = k % 6 == 3
    d[tiny, (k[tiny] // 6) % 3] *= np.array([1e-6, 1e-30, 0.0])[(k[tiny] // 18) % 3]
    scaled = k % 6 == 4
    d[scaled] *= 10.0 ** rng.uniform(-3, 3, size=(int(scaled.sum()), 1))
    d = d.astype(np.float32)
    ref = ref_svo.trace(o, d, 1024.0)
    try:
        for dda, cross, node, burst, chunk, refill, small in ((0, 0, 1, 8, 32, 10, 4), (1, 0, 1, 8, 0, 1, 0),
                                                              (1, 1, 2, 2, 256, 64, 4), (1, 32, 3, 8, 32, 10, 0),
                                                              (1, 40, 4, 16, 64, 4, 64), (1, 64, 3, 32, 256, 1, 4)):
            ctx.set_tuning(trace_dda=dda, trace_cross=cross, trace_node=node, trace_burst=burst, trace_chunk=chunk,
                           trace_refill=refill, trace_small=small)
            say(f"depth-11 trace, 24000 stress rays, trace_dda {dda} trace_cross {cross} trace_node {node} "
                f"trace_burst {burst} trace_chunk {chunk} trace_refill {refill} trace_small {small}")
            assert_same_hits(ctx.svo_trace(o, d, 1024.0), ref, f"dda {dda} cross {cross} node {node} chunk {chunk}")
    finally:
        ctx.set_tuning(trace_dda=1, trace_cross=40, trace_node=3, trace_burst=8, trace_chunk=32, trace_refill=5,
                       trace_small=8)
    assert 0.2 < ref["hit"].mean() < 0.95
