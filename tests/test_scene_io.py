"""Scene files (tgx_scene_save / tgx_scene_file_counts / tgx_scene_load):
byte-exact round trips of generated scenes, materials optional, and
rejection of corrupted, truncated and foreign files. Host only."""
from __future__ import annotations

import numpy as np
import pytest

import oracle as O
import tg_amd as T

FIELDS = ("objects", "cluster_pointers", "cluster_idx_to_object", "bits", "materials", "luts")


@pytest.mark.parametrize("params", [T.single_object_params(4),
                                    T.noise_objects_params(seed=5, n_objects=30, max_extent=6, world=128.0, n_luts=3)])
def test_round_trip_is_byte_exact(tmp_path, params):
    sc = T.generate_scene(params)
    path = tmp_path / "scene.tgxs"
    T.save_scene(path, sc)
    got = T.load_scene(path)
    for f in FIELDS:
        assert getattr(got, f).tobytes() == getattr(sc, f).tobytes(), f
    # header + arrays + checksum, nothing else
    n = sc.n_clusters
    expect = 8 + 24 + 96 * sc.objects.size + 4 * sc.cluster_pointers.size + 4 * n + 64 * n + 512 * n \
        + 16 * sc.luts.size + 8
    assert path.stat().st_size == expect
    # the loaded scene renders the same visibility (oracle)
    cam = T.generate_camera(params, 0, 64, 48)
    a, _ = O.visibility_culled(O.OracleScene(sc), cam, 64, 48)
    b, _ = O.visibility_culled(O.OracleScene(got), cam, 64, 48)
    assert np.array_equal(a, b)


def test_scene_without_materials_loads_zeros(tmp_path):
    sc = T.generate_scene(T.single_object_params(2))
    lib = T.load_library()
    d = sc.desc()
    d.materials = None
    path = tmp_path / "nomat.tgxs"
    assert lib.tgx_scene_save(str(path).encode(), d) == 0
    got = T.load_scene(path)
    assert not got.materials.any() and got.bits.tobytes() == sc.bits.tobytes()


def test_corrupt_truncated_and_foreign_files_are_rejected(tmp_path):
    sc = T.generate_scene(T.single_object_params(3))
    path = tmp_path / "s.tgxs"
    T.save_scene(path, sc)
    raw = bytearray(path.read_bytes())
    flipped = bytearray(raw)
    flipped[len(raw) // 2] ^= 0x10
    (tmp_path / "flip.tgxs").write_bytes(bytes(flipped))
    (tmp_path / "cut.tgxs").write_bytes(bytes(raw[:-9]))
    (tmp_path / "long.tgxs").write_bytes(bytes(raw) + b"\0")
    foreign = bytearray(raw)
    foreign[0:8] = b"NOTASCN0"
    (tmp_path / "foreign.tgxs").write_bytes(bytes(foreign))
    for name in ("flip", "cut", "long", "foreign"):
        with pytest.raises(T.TgxError) as e:
            T.load_scene(tmp_path / f"{name}.tgxs")
        assert e.value.code == 1, name  # TGX_ERR_INVALID_ARG
    with pytest.raises(T.TgxError) as e:
        T.load_scene(tmp_path / "missing.tgxs")
    assert e.value.code == 8  # TGX_ERR_IO


def test_load_refuses_counts_that_differ_from_the_header(tmp_path):
    """A file replaced between tgx_scene_file_counts and tgx_scene_load (here:
    counts of a smaller scene) is refused before any caller array is written."""
    small = T.generate_scene(T.single_object_params(1))
    big = T.generate_scene(T.noise_objects_params(seed=5, n_objects=30, max_extent=6, world=128.0, n_luts=3))
    path = tmp_path / "s.tgxs"
    T.save_scene(path, big)
    lib = T.load_library()
    counts = np.zeros(1, T.SCENE_COUNTS_DTYPE)
    counts["n_objects"], counts["n_cluster_pointers"] = small.objects.size, small.n_cluster_pointers
    counts["n_clusters"], counts["n_luts"] = small.n_clusters, small.luts.size // 256
    guard = 4096
    arrays = [np.zeros(small.objects.size + guard, T.OBJECT_DTYPE),
              np.zeros(small.n_cluster_pointers + guard, np.uint32), np.zeros(small.n_clusters + guard, np.uint32),
              np.zeros(16 * small.n_clusters + guard, np.uint32), np.zeros(512 * small.n_clusters + guard, np.uint8),
              np.zeros(small.luts.size + guard, T.LUT_DTYPE)]
    rc = lib.tgx_scene_load(str(path).encode(), counts.ctypes.data, *[a.ctypes.data for a in arrays])
    assert rc == 1  # TGX_ERR_INVALID_ARG
    assert not any(a.view(np.uint8).any() for a in arrays)
