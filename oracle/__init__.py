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
