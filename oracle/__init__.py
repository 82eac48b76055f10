"""CPU oracle of the voxel ray-tracing hot path — TEST INFRASTRUCTURE ONLY.

ctypes view of oracle/liboracle.so (built by ``make -C oracle``), a strict
IEEE fp32 C restatement of tg's algorithms. Only tests/,
__graft_entry__.smoke() and bench.py's CPU-baseline leg may import this
package, and only as the checker. The product (tg_amd, libtgx.so) never
uses it.
"""
from __future__ import annotations

import ctypes
from pathlib import Path

import numpy as np

ORACLE_DIR = Path(__file__).resolve().parent
LIB_PATH = ORACLE_DIR / "liboracle.so"


class _OScene(ctypes.Structure):
    _fields_ = [
        ("objects", ctypes.c_void_p), ("n_objects", ctypes.c_uint32),
        ("cluster_pointers", ctypes.c_void_p), ("n_cluster_pointers", ctypes.c_uint32),
        ("cluster_idx_to_object", ctypes.c_void_p), ("n_clusters", ctypes.c_uint32),
        ("bits", ctypes.c_void_p), ("materials", ctypes.c_void_p),
        ("luts", ctypes.c_void_p), ("n_luts", ctypes.c_uint32),
    ]


_LIB = None


def load() -> ctypes.CDLL:
    global _LIB
    if _LIB is not None:
        return _LIB
    if not LIB_PATH.exists():
        raise RuntimeError(f"{LIB_PATH} is missing: run `make -C oracle`")
    lib = ctypes.CDLL(str(LIB_PATH))
    vp, u32, u64, i32 = ctypes.c_void_p, ctypes.c_uint32, ctypes.c_uint64, ctypes.c_int
    fp = ctypes.POINTER(ctypes.c_float)
    sigs = {
        "tgxo_vis_pair": (u64, [vp, vp, u32, u32, u32, u32, u32]),
        "tgxo_visibility_brute": (None, [vp, vp, u32, u32, vp]),
        "tgxo_visibility_culled": (u64, [vp, vp, u32, u32, u32, i32, vp]),
        "tgxo_cluster_rect": (i32, [vp, vp, u32, u32, u32, vp]),
        "tgxo_visibility_alg_bytes": (u64, [vp, vp, u32, u32]),
        "tgxo_intersect_ray_aabb": (i32, [vp, vp, vp, vp, fp, fp]),
        "tgxo_m4_mul": (None, [vp, vp, vp]),
        "tgxo_m4_inverse": (None, [vp, vp]),
        "tgxo_world_to_cluster": (None, [vp, u32, vp]),
        "tgxo_svo_build": (vp, [vp, vp, u32, u32]),
        "tgxo_svo_free": (None, [vp]),
        "tgxo_svo_get_info": (None, [vp, vp]),
        "tgxo_svo_export": (i32, [vp, vp, vp, vp, vp, vp, vp]),
        "tgxo_svo_traverse": (None, [vp, vp, vp, ctypes.c_float, vp]),
        "tgxo_svo_trace": (None, [vp, vp, vp, u32, ctypes.c_float, vp]),
        "tgxo_aabb_obb_ignore_contact": (i32, [vp, vp, vp]),
        "tgxo_svo_node_key": (u64, [u32, vp]),
        "tgxo_gi": (None, [vp, vp, u32, u32, vp, vp, u32, vp, vp, vp]),
        "tgxo_shade": (None, [vp, vp, u32, u32, vp, vp, vp]),
        "tgxo_shade_pixel": (None, [vp, vp, u32, u32, u32, u32, u64, vp, vp]),
        "tgxo_gi_surface": (i32, [vp, vp, u32, u32, u32, u32, u64, vp, vp, vp]),
        "tgxo_gi_surface_count": (u64, [vp, vp, u32, u32, vp, i32]),
        "tgxo_gi_seed": (u32, [u32, u32, u32, u32]),
        "tgxo_gi_direction": (None, [vp, u32, vp]),
        "tgxo_gi_alg_bytes": (i32, [vp, vp, u32, u32, vp, vp, u32, vp]),
    }
    for name, (res, args) in sigs.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _LIB = lib
    return lib


def _p(a):
    if a is None:
        return None
    assert a.flags["C_CONTIGUOUS"]
    return a.ctypes.data


class OracleScene:
    """Holds a scene's arrays and the C view of them."""

    def __init__(self, scene):
        self.scene = scene
        s = scene
        self.c = _OScene(_p(s.objects), s.objects.size, _p(s.cluster_pointers), s.cluster_pointers.size,
                         _p(s.cluster_idx_to_object), s.cluster_idx_to_object.size, _p(s.bits),
                         _p(s.materials), _p(s.luts), s.luts.size // 256)

    @property
    def ref(self):
        return ctypes.byref(self.c)


def visibility_brute(oscene: OracleScene, cam: np.ndarray, w: int, h: int) -> np.ndarray:
    out = np.empty((h, w), np.uint64)
    load().tgxo_visibility_brute(oscene.ref, _p(cam), w, h, _p(out))
    return out


def visibility_culled(oscene: OracleScene, cam: np.ndarray, w: int, h: int, row_step: int = 1,
                      n_threads: int = 0) -> tuple[np.ndarray, int]:
    out = np.full((h, w), 0xFFFFFFFFFFFFFFFF, np.uint64)
    pairs = load().tgxo_visibility_culled(oscene.ref, _p(cam), w, h, row_step, n_threads, _p(out))
    return out, int(pairs)


def gi_surface_count(oscene: OracleScene, cam: np.ndarray, w: int, h: int, vis: np.ndarray,
                     n_threads: int = 0) -> int:
    """Pixels of one camera's visibility words that have a primary surface
    (the visible pixels a GI frame lists)."""
    v = np.ascontiguousarray(vis, np.uint64)
    assert v.size == w * h
    return int(load().tgxo_gi_surface_count(oscene.ref, _p(cam), w, h, _p(v), n_threads))


def shade(oscene: OracleScene, cam: np.ndarray, w: int, h: int, vis: np.ndarray, params: np.ndarray) -> np.ndarray:
    """Shading pass of one camera: RGBA float32 [h, w, 4]."""
    vis = np.ascontiguousarray(vis, np.uint64)
    out = np.zeros((h, w, 4), np.float32)
    load().tgxo_shade(oscene.ref, _p(cam), w, h, _p(vis), _p(params), _p(out))
    return out


def shade_pixel(oscene: OracleScene, cam: np.ndarray, w: int, h: int, x: int, y: int, word: int,
                params: np.ndarray) -> np.ndarray:
    out = np.zeros(4, np.float32)
    load().tgxo_shade_pixel(oscene.ref, _p(cam), w, h, x, y, int(word), _p(params), _p(out))
    return out


def vis_pair(oscene: OracleScene, cam: np.ndarray, w: int, h: int, x: int, y: int, cptr: int) -> int:
    return int(load().tgxo_vis_pair(oscene.ref, _p(cam), w, h, x, y, cptr))


def cluster_rect(oscene: OracleScene, cam: np.ndarray, w: int, h: int, cptr: int):
    r = np.zeros(4, np.uint32)
    ok = load().tgxo_cluster_rect(oscene.ref, _p(cam), w, h, cptr, _p(r))
    return (bool(ok), tuple(int(v) for v in r))


def visibility_alg_bytes(oscene: OracleScene, cam: np.ndarray, w: int, h: int) -> int:
    return int(load().tgxo_visibility_alg_bytes(oscene.ref, _p(cam), w, h))


def intersect_ray_aabb(o, d, mn, mx) -> tuple[bool, float, float]:
    a = [np.ascontiguousarray(v, np.float32) for v in (o, d, mn, mx)]
    e, x = ctypes.c_float(), ctypes.c_float()
    hit = load().tgxo_intersect_ray_aabb(_p(a[0]), _p(a[1]), _p(a[2]), _p(a[3]), ctypes.byref(e), ctypes.byref(x))
    return bool(hit), e.value, x.value


def m4_mul(a: np.ndarray, b: np.ndarray) -> np.ndarray:
    a = np.ascontiguousarray(a, np.float32); b = np.ascontiguousarray(b, np.float32)
    out = np.zeros(16, np.float32)
    load().tgxo_m4_mul(_p(a), _p(b), _p(out))
    return out


def m4_inverse(a: np.ndarray) -> np.ndarray:
    a = np.ascontiguousarray(a, np.float32)
    out = np.zeros(16, np.float32)
    load().tgxo_m4_inverse(_p(a), _p(out))
    return out


def world_to_cluster(obj: np.ndarray, rel_cptr: int) -> np.ndarray:
    obj = np.ascontiguousarray(obj)
    out = np.zeros(16, np.float32)
    load().tgxo_world_to_cluster(_p(obj), rel_cptr, _p(out))
    return out


# --------------------------------------------------------------------------
# Sparse voxel octree
# --------------------------------------------------------------------------

SVO_INFO_DTYPE = np.dtype([("min", "<f4", (3,)), ("depth", "<u4"), ("n_nodes", "<u4"), ("n_leaves", "<u4"),
                           ("n_list_entries", "<u4"), ("flags", "<u4")])
SVO_NODE_DTYPE = np.dtype([("level", "<u4"), ("coord", "<u4", (3,)), ("valid_mask", "<u4"), ("leaf_mask", "<u4"),
                           ("first_child", "<u4"), ("leaf_idx", "<u4")])
SVO_HIT_DTYPE = np.dtype([("hit", "<u4"), ("distance", "<f4"), ("block", "<u4", (3,)), ("voxel", "<u4"),
                          ("normal", "<i4", (3,)), ("steps", "<u4")])
SVO_FLAG_REFERENCE_CORNER_QUIRK = 1


class SvoExport:
    """Canonical SVO arrays (same layout from the oracle and from the GPU)."""

    def __init__(self, info, nodes, bits, materials, leaf_n, list_offsets, lst):
        self.info, self.nodes, self.bits, self.materials = info, nodes, bits, materials
        self.leaf_n, self.list_offsets, self.list = leaf_n, list_offsets, lst

    @staticmethod
    def alloc(info) -> "SvoExport":
        nl = int(info["n_leaves"])
        return SvoExport(info, np.zeros(int(info["n_nodes"]), SVO_NODE_DTYPE), np.zeros((nl, 1024), np.uint32),
                         np.zeros((nl, 32768), np.uint16), np.zeros(nl, np.uint32), np.zeros(nl + 1, np.uint32),
                         np.zeros(max(1, int(info["n_list_entries"])), np.uint32))

    def leaf_coords(self) -> np.ndarray:
        leaves = self.nodes[self.nodes["leaf_idx"] != 0xFFFFFFFF]
        order = np.argsort(leaves["leaf_idx"])
        return leaves["coord"][order]


class OracleSvo:
    def __init__(self, oscene: OracleScene, mn, depth: int, flags: int = 0):
        self.oscene = oscene  # keep the scene arrays alive
        m = np.ascontiguousarray(mn, np.float32)
        self.h = load().tgxo_svo_build(oscene.ref, _p(m), depth, flags)
        if not self.h:
            raise RuntimeError("tgxo_svo_build failed")

    def __del__(self):
        if getattr(self, "h", None):
            load().tgxo_svo_free(self.h)
            self.h = None

    def info(self):
        i = np.zeros(1, SVO_INFO_DTYPE)
        load().tgxo_svo_get_info(self.h, _p(i))
        return i[0]

    def export(self) -> SvoExport:
        e = SvoExport.alloc(self.info())
        load().tgxo_svo_export(self.h, _p(e.nodes), _p(e.bits), _p(e.materials), _p(e.leaf_n),
                               _p(e.list_offsets), _p(e.list))
        return e

    def gi(self, cam: np.ndarray, w: int, h: int, vis: np.ndarray, params: np.ndarray, row_step: int = 1):
        """(radiance[h, w, 3], irradiance[h, w, 3], counts[4]) of one camera."""
        vis = np.ascontiguousarray(vis, np.uint64)
        rad = np.zeros((h, w, 3), np.float32)
        irr = np.zeros((h, w, 3), np.float32)
        cnt = np.zeros(4, np.uint64)
        load().tgxo_gi(self.h, _p(cam), w, h, _p(vis), _p(params), row_step, _p(rad), _p(irr), _p(cnt))
        return rad, irr, cnt

    def gi_alg_bytes(self, cam: np.ndarray, w: int, h: int, vis: np.ndarray, params: np.ndarray,
                     with_irradiance: bool = False) -> dict:
        """Algorithmic bytes of the GI frame (tgxo_gi_alg_bytes): unique 64-B
        lines the oracle's execution reads, per part, plus bytes written."""
        vis = np.ascontiguousarray(vis, np.uint64)
        out = np.zeros(16, np.uint64)
        if load().tgxo_gi_alg_bytes(self.h, _p(cam), w, h, _p(vis), _p(params), int(with_irradiance), _p(out)):
            raise RuntimeError("tgxo_gi_alg_bytes: allocation failed")
        names = ("nodes", "leaf_data", "voxel_pool", "leaf_materials", "luts_hits", "cluster_pointers",
                 "cluster_to_object", "objects", "scene_materials", "luts_surfaces", "visibility")
        return {"surface": int(out[0]), "trace": int(out[1]), "written": int(out[2]), "total": int(out[3]),
                "lines": {k: int(v) for k, v in zip(names, out[4:15])}}

    def trace(self, origins: np.ndarray, dirs: np.ndarray, max_distance: float = 3.0e38) -> np.ndarray:
        o = np.ascontiguousarray(origins, np.float32).reshape(-1, 3)
        d = np.ascontiguousarray(dirs, np.float32).reshape(-1, 3)
        out = np.zeros(o.shape[0], SVO_HIT_DTYPE)
        load().tgxo_svo_trace(self.h, _p(o), _p(d), o.shape[0], max_distance, _p(out))
        return out


def aabb_obb_ignore_contact(mn, mx, corners) -> bool:
    a = [np.ascontiguousarray(v, np.float32) for v in (mn, mx, corners)]
    return bool(load().tgxo_aabb_obb_ignore_contact(_p(a[0]), _p(a[1]), _p(a[2])))


def svo_node_key(level: int, coord) -> int:
    c = np.ascontiguousarray(coord, np.uint32)
    return int(load().tgxo_svo_node_key(level, _p(c)))
