"""tg_amd — MI355X-native voxel ray-tracing hot path of tg.

Python view of the C ABI in include/tgx.h (libtgx.so, built in-tree by
``python -m tg_amd.build``). The GPU work happens in libtgx.so; this module
only moves numpy arrays across the boundary. There is no CPU fallback: if
libtgx.so is missing, importing the bindings raises.

Byte layouts (numpy dtypes below) are tg's std430 records
(buffers.inc:6-13 objects, buffers.inc:30-39 camera).
"""
from __future__ import annotations

import ctypes
from dataclasses import dataclass
from pathlib import Path

import numpy as np

PKG_DIR = Path(__file__).resolve().parent
LIB_PATH = PKG_DIR / "libtgx.so"

OBJECT_DTYPE = np.dtype([
    ("n_cluster_pointers_per_dim", "<u4", (3,)),
    ("first_cluster_pointer", "<u4"),
    ("translation", "<f4", (3,)),
    ("lut_idx", "<u4"),
    ("rotation", "<f4", (16,)),
])
CAMERA_DTYPE = np.dtype([
    ("camera", "<f4", (4,)),
    ("ray_bl", "<f4", (4,)),
    ("ray_br", "<f4", (4,)),
    ("ray_tr", "<f4", (4,)),
    ("ray_tl", "<f4", (4,)),
    ("near_plane", "<f4"),
    ("far_plane", "<f4"),
])
LUT_DTYPE = np.dtype([("albedo", "<f4", (3,)), ("roughness", "<f4")])
assert OBJECT_DTYPE.itemsize == 96 and CAMERA_DTYPE.itemsize == 88 and LUT_DTYPE.itemsize == 16

SVO_INFO_DTYPE = np.dtype([("min", "<f4", (3,)), ("depth", "<u4"), ("n_nodes", "<u4"), ("n_leaves", "<u4"),
                           ("n_list_entries", "<u4"), ("flags", "<u4")])
SVO_NODE_DTYPE = np.dtype([("level", "<u4"), ("coord", "<u4", (3,)), ("valid_mask", "<u4"), ("leaf_mask", "<u4"),
                           ("first_child", "<u4"), ("leaf_idx", "<u4")])
SVO_HIT_DTYPE = np.dtype([("hit", "<u4"), ("distance", "<f4"), ("block", "<u4", (3,)), ("voxel", "<u4"),
                          ("normal", "<i4", (3,)), ("steps", "<u4")])
SVO_FLAG_REFERENCE_CORNER_QUIRK = 1
GI_PARAMS_DTYPE = np.dtype([("sun_dir", "<f4", (3,)), ("sun_radiance", "<f4"), ("spp", "<u4"),
                            ("max_distance", "<f4"), ("frame", "<u4"), ("pad", "<u4")])
GI_STATS_DTYPE = np.dtype([("visible_pixels", "<u8"), ("bounce_rays", "<u8"), ("bounce_hits", "<u8"),
                           ("shadow_rays", "<u8")])


def gi_params(sun_dir=(0.3, 0.8, 0.2), sun_radiance: float = 1.0, spp: int = 4, max_distance: float = 1024.0,
              frame: int = 0) -> np.ndarray:
    p = np.zeros(1, GI_PARAMS_DTYPE)
    d = np.asarray(sun_dir, np.float64)
    p["sun_dir"] = (d / np.linalg.norm(d)).astype(np.float32)
    p["sun_radiance"] = sun_radiance
    p["spp"] = spp
    p["max_distance"] = max_distance
    p["frame"] = frame
    return p


def seeded_sun(seed: int) -> np.ndarray:
    """Fixed seeded sun direction: azimuth U[0, 2pi), elevation U[25, 65] deg."""
    rng = np.random.default_rng(np.random.SeedSequence([seed, 0x5E1]))
    az, el = rng.uniform(0, 2 * np.pi), np.radians(rng.uniform(25, 65))
    return np.array([np.cos(el) * np.cos(az), np.sin(el), np.cos(el) * np.sin(az)])

TUNING_DTYPE = np.dtype([("trace_sort", "<u4"), ("trace_refill", "<u4"), ("trace_blocks_per_cu", "<u4"),
                         ("trace_vote", "<u4"), ("trace_regroup", "<u4"), ("trace_skip", "<u4"),
                         ("vis_setup_blocks", "<u4"), ("trace_dda", "<u4"),
                         ("vis_tile", "<u4"), ("trace_cross", "<u4"),
                         ("trace_node", "<u4"), ("trace_burst", "<u4"), ("trace_chunk", "<u4"),
                         ("trace_small", "<u4")])

NRC_N_PARAMS = 18624
NRC_CONFIG_DTYPE = np.dtype([("origin", "<f4", (3,)), ("scale", "<f4"), ("learning_rate", "<f4"),
                             ("beta1", "<f4"), ("beta2", "<f4"), ("epsilon", "<f4"), ("seed", "<u4"),
                             ("train_samples", "<u4"), ("train_batch", "<u4"), ("train_spp", "<u4"),
                             ("fallback_threshold", "<f4"), ("fallback_min_samples", "<u4"),
                             ("cell_size", "<u4"), ("flags", "<u4"), ("ema_decay", "<f4")])
NRC_STATS_DTYPE = np.dtype([("visible_pixels", "<u8"), ("fallback_pixels", "<u8"), ("train_samples", "<u8"),
                            ("train_rays", "<u8"), ("fallback_rays", "<u8"), ("adam_steps", "<u4"),
                            ("train_loss", "<f4")])
SHADE_PARAMS_DTYPE = np.dtype([("view", "<u4"), ("light_dir", "<f4", (3,)), ("light_radiance", "<f4"),
                               ("ambient", "<f4"), ("metallic", "<f4"), ("roughness", "<f4"), ("miss", "<f4", (4,))])
SHADE_LIT, SHADE_OBJECT_INDEX, SHADE_DEPTH, SHADE_CLUSTER_INDEX, SHADE_VOXEL_INDEX = 0, 1, 2, 3, 4
SHADE_BLOCKS, SHADE_COLOR_LUT_INDEX, SHADE_COLOR, SHADE_NORMAL, SHADE_SHADING = 5, 6, 7, 8, 9


def shade_params(view: int = SHADE_LIT, light_dir=(0.0, 0.8, 0.3), light_radiance: float = 3.0,
                 ambient: float = 0.1, metallic: float = 0.1, roughness: float = 0.8,
                 miss=(1.0, 0.0, 1.0, 1.0)) -> np.ndarray:
    """Shading parameters; the defaults are tg's constants (shading.frag:308-322)."""
    p = np.zeros(1, SHADE_PARAMS_DTYPE)
    p["view"] = view
    p["light_dir"] = light_dir
    p["light_radiance"] = light_radiance
    p["ambient"] = ambient
    p["metallic"] = metallic
    p["roughness"] = roughness
    p["miss"] = miss
    return p


PICK_DTYPE = np.dtype([("hit", "<u4"), ("depth", "<f4"), ("cluster_pointer", "<u4"), ("cluster_idx", "<u4"),
                       ("object_idx", "<u4"), ("voxel", "<u4"), ("position", "<f4", (3,))])
SCENE_COUNTS_DTYPE = np.dtype([("n_objects", "<u4"), ("n_cluster_pointers", "<u4"), ("n_clusters", "<u4"),
                               ("n_luts", "<u4"), ("has_materials", "<u4")])

FRAME_NEURAL = 1
FRAME_STATS_DTYPE = np.dtype([("primary_rays", "<u8"), ("visible_pixels", "<u8"), ("bounce_rays", "<u8"),
                              ("shadow_rays", "<u8"), ("bounce_hits", "<u8"), ("train_rays", "<u8"),
                              ("fallback_pixels", "<u8")])

NRC_FLAG_NO_FALLBACK = 1
NRC_FLAG_ALL_FALLBACK = 2
NRC_FLAG_NO_TRAINING = 4
NRC_FLAG_FALLBACK_VS_TRACER = 8
NRC_FLAG_JOINT_TRAINING = 16


def nrc_config(origin=(0.0, 0.0, 0.0), scale: float = 1.0 / 512, learning_rate: float = 1e-3, beta1: float = 0.9,
               beta2: float = 0.99, epsilon: float = 1e-8, seed: int = 1, train_samples: int = 1 << 16,
               train_batch: int = 1 << 12, train_spp: int = 4, fallback_threshold: float = 1.0,
               fallback_min_samples: int = 1, cell_size: int = 8, flags: int = 0,
               ema_decay: float = 0.0) -> np.ndarray:
    c = np.zeros(1, NRC_CONFIG_DTYPE)
    c["origin"] = origin
    for k, v in (("scale", scale), ("learning_rate", learning_rate), ("beta1", beta1), ("beta2", beta2),
                 ("epsilon", epsilon), ("seed", seed), ("train_samples", train_samples),
                 ("train_batch", train_batch), ("train_spp", train_spp), ("fallback_threshold", fallback_threshold),
                 ("fallback_min_samples", fallback_min_samples), ("cell_size", cell_size), ("flags", flags),
                 ("ema_decay", ema_decay)):
        c[k] = v
    return c


VISIBILITY_EMPTY = np.uint64(0xFFFFFFFFFFFFFFFF)

GEN_SINGLE_BERNOULLI = 0
GEN_NOISE_OBJECTS = 1

STATUS = {0: "TGX_OK", 1: "TGX_ERR_INVALID_ARG", 2: "TGX_ERR_CAPACITY", 3: "TGX_ERR_HIP",
          4: "TGX_ERR_NO_SCENE", 5: "TGX_ERR_NO_DEVICE", 6: "TGX_ERR_OUT_OF_MEMORY",
          7: "TGX_ERR_NOT_BUILT", 8: "TGX_ERR_IO"}


class TgxError(RuntimeError):
    def __init__(self, code: int, message: str = ""):
        self.code = code
        super().__init__(f"{STATUS.get(code, code)}: {message}")


class _SceneDesc(ctypes.Structure):
    _fields_ = [
        ("objects", ctypes.c_void_p), ("n_objects", ctypes.c_uint32),
        ("cluster_pointers", ctypes.c_void_p), ("n_cluster_pointers", ctypes.c_uint32),
        ("cluster_idx_to_object", ctypes.c_void_p), ("n_clusters", ctypes.c_uint32),
        ("bits", ctypes.c_void_p), ("materials", ctypes.c_void_p),
        ("luts", ctypes.c_void_p), ("n_luts", ctypes.c_uint32),
    ]


class _GenParams(ctypes.Structure):
    _fields_ = [("kind", ctypes.c_uint32), ("seed", ctypes.c_uint64), ("n_objects", ctypes.c_uint32),
                ("max_extent", ctypes.c_uint32), ("world", ctypes.c_float), ("n_luts", ctypes.c_uint32)]


class _GenSizes(ctypes.Structure):
    _fields_ = [("n_objects", ctypes.c_uint32), ("n_cluster_pointers", ctypes.c_uint32),
                ("n_clusters", ctypes.c_uint32), ("n_luts", ctypes.c_uint32)]


class KernelStat(ctypes.Structure):
    _fields_ = [("name", ctypes.c_char * 48), ("launches", ctypes.c_uint64), ("total_ms", ctypes.c_double)]


class VisibilityStats(ctypes.Structure):
    _fields_ = [("clusters_total", ctypes.c_uint64), ("clusters_binned", ctypes.c_uint64),
                ("tile_entries", ctypes.c_uint64), ("tiles", ctypes.c_uint32), ("tile_size", ctypes.c_uint32)]


_LIB: ctypes.CDLL | None = None


def load_library() -> ctypes.CDLL:
    """Loads libtgx.so; raises if it has not been built."""
    global _LIB
    if _LIB is not None:
        return _LIB
    if not LIB_PATH.exists():
        raise RuntimeError(f"{LIB_PATH} is missing: run `python -m tg_amd.build` (there is no CPU fallback)")
    lib = ctypes.CDLL(str(LIB_PATH))
    vp, u32, u64, i32 = ctypes.c_void_p, ctypes.c_uint32, ctypes.c_uint64, ctypes.c_int
    sigs = {
        "tgx_create": (i32, [i32, ctypes.POINTER(vp)]),
        "tgx_destroy": (i32, [vp]),
        "tgx_last_error": (ctypes.c_char_p, [vp]),
        "tgx_synchronize": (i32, [vp]),
        "tgx_scene_upload": (i32, [vp, ctypes.POINTER(_SceneDesc)]),
        "tgx_visibility": (i32, [vp, vp, u32, u32, u32, vp]),
        "tgx_visibility_device": (i32, [vp, vp, u32, u32, u32, vp]),
        "tgx_visibility_device_noalloc": (i32, [vp, vp, u32, u32, u32, vp]),
        "tgx_nrc_save": (i32, [vp, ctypes.c_char_p]),
        "tgx_svo_device_bytes": (i32, [vp, ctypes.POINTER(u64), ctypes.POINTER(u64)]),
        "tgx_nrc_load": (i32, [vp, ctypes.c_char_p]),
        "tgx_visibility_clear_device": (i32, [vp, vp, u64]),
        "tgx_visibility_decode": (i32, [u64, ctypes.POINTER(ctypes.c_float), ctypes.POINTER(u32), ctypes.POINTER(u32)]),
        "tgx_visibility_last_stats": (i32, [vp, ctypes.POINTER(VisibilityStats)]),
        "tgx_visibility_set_list_capacity": (i32, [vp, u32]),
        "tgx_device_alloc": (i32, [vp, ctypes.c_size_t, ctypes.POINTER(vp)]),
        "tgx_device_free": (i32, [vp, vp]),
        "tgx_memcpy_d2h": (i32, [vp, vp, vp, ctypes.c_size_t]),
        "tgx_memcpy_h2d": (i32, [vp, vp, vp, ctypes.c_size_t]),
        "tgx_memcpy_d2d": (i32, [vp, vp, vp, ctypes.c_size_t]),
        "tgx_gen_scene_device": (i32, [vp, ctypes.POINTER(_GenParams)]),
        "tgx_scene_export": (i32, [vp, vp, vp, vp, vp, vp, vp]),
        "tgx_tuning_get": (i32, [vp, vp]),
        "tgx_tuning_set": (i32, [vp, vp]),
        "tgx_profile_enable": (i32, [vp, i32]),
        "tgx_profile_reset": (i32, [vp]),
        "tgx_profile_get": (i32, [vp, ctypes.POINTER(KernelStat), u32, ctypes.POINTER(u32)]),
        "tgx_gen_scene_sizes": (i32, [ctypes.POINTER(_GenParams), ctypes.POINTER(_GenSizes)]),
        "tgx_gen_scene_fill": (i32, [ctypes.POINTER(_GenParams), vp, vp, vp, vp, vp, vp]),
        "tgx_gen_camera": (i32, [ctypes.POINTER(_GenParams), u32, u32, u32, vp]),
        "tgx_camera_make": (i32, [vp, ctypes.c_float, ctypes.c_float, ctypes.c_float, ctypes.c_float,
                                  ctypes.c_float, ctypes.c_float, ctypes.c_float, vp]),
        "tgx_gen_stream_seed": (u64, [u64, u64]),
        "tgx_svo_build": (i32, [vp, vp, u32, u32]),
        "tgx_svo_update": (i32, [vp, vp, vp]),
        "tgx_lut_set": (i32, [vp, u32, u32, vp]),
        "tgx_svo_get_info": (i32, [vp, vp]),
        "tgx_svo_export": (i32, [vp, vp, vp, vp, vp, vp, vp]),
        "tgx_svo_save": (i32, [vp, ctypes.c_char_p]),
        "tgx_svo_load": (i32, [vp, ctypes.c_char_p]),
        "tgx_svo_trace": (i32, [vp, vp, vp, u32, ctypes.c_float, vp]),
        "tgx_gi_device": (i32, [vp, vp, u32, u32, vp, vp, vp, vp]),
        "tgx_gi_last_stats": (i32, [vp, vp]),
        "tgx_frame_device": (i32, [vp, vp, u32, u32, vp, vp, vp, vp]),
        "tgx_nrc_create": (i32, [vp, vp]),
        "tgx_nrc_get_params": (i32, [vp, vp]),
        "tgx_nrc_get_inference_params": (i32, [vp, vp]),
        "tgx_nrc_set_params": (i32, [vp, vp]),
        "tgx_nrc_infer": (i32, [vp, vp, u32, vp]),
        "tgx_nrc_train_batch": (i32, [vp, vp, vp, u32, vp, vp]),
        "tgx_nrc_frame_device": (i32, [vp, vp, u32, u32, vp, vp, vp, vp]),
        "tgx_nrc_last_stats": (i32, [vp, vp]),
        "tgx_frame_graph_create": (i32, [vp, vp, u32, u32, u32, vp, u32, vp, vp, vp, ctypes.POINTER(vp)]),
        "tgx_frame_batch": (i32, [vp, vp, u32, u32, u32, vp, u32, vp, vp, vp, vp]),
        "tgx_shade_device": (i32, [vp, vp, u32, u32, vp, vp, vp]),
        "tgx_pick": (i32, [vp, vp, vp, u32, u32, vp, u32, vp]),
        "tgx_svo_visibility_device": (i32, [vp, vp, u32, u32, vp]),
        "tgx_scene_get_counts": (i32, [vp, vp]),
        "tgx_scene_reserve": (i32, [vp, u32, u32, vp, u32]),
        "tgx_object_create": (i32, [vp, vp, vp, vp, u32, vp, vp, ctypes.POINTER(u32)]),
        "tgx_object_destroy": (i32, [vp, u32]),
        "tgx_scene_save": (i32, [ctypes.c_char_p, ctypes.POINTER(_SceneDesc)]),
        "tgx_scene_file_counts": (i32, [ctypes.c_char_p, vp]),
        "tgx_scene_load": (i32, [ctypes.c_char_p, vp, vp, vp, vp, vp, vp, vp]),
        "tgx_frame_graph_launch": (i32, [vp, vp]),
        "tgx_frame_graph_stats": (i32, [vp, vp, vp]),
        "tgx_frame_graph_destroy": (i32, [vp, vp]),
        "tgx_autotune": (i32, [vp, vp, u32, u32, u32, vp, u32, vp, vp, vp, u32, vp]),
    }
    for name, (res, args) in sigs.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _LIB = lib
    return lib


def exported_symbols() -> list[str]:
    """Names of the C-ABI functions the bindings declare."""
    return [n for n in dir(load_library()) if n.startswith("tgx_")]


def _ptr(a: np.ndarray | None) -> int | None:
    if a is None:
        return None
    assert a.flags["C_CONTIGUOUS"], "arrays crossing the C ABI must be C-contiguous"
    return a.ctypes.data


# --------------------------------------------------------------------------
# Scenes
# --------------------------------------------------------------------------

@dataclass
class Scene:
    """tg's object/cluster data model as numpy arrays (include/tgx.h)."""
    objects: np.ndarray                 # OBJECT_DTYPE[n_objects]
    cluster_pointers: np.ndarray        # u32[n_cluster_pointers]: cptr -> cidx
    cluster_idx_to_object: np.ndarray   # u32[n_clusters]
    bits: np.ndarray                    # u32[16 * n_clusters]
    materials: np.ndarray               # u8[512 * n_clusters]
    luts: np.ndarray                    # LUT_DTYPE[256 * n_luts]

    @property
    def n_clusters(self) -> int:
        return int(self.cluster_idx_to_object.size)

    @property
    def n_cluster_pointers(self) -> int:
        return int(self.cluster_pointers.size)

    def desc(self) -> _SceneDesc:
        return _SceneDesc(
            _ptr(self.objects), self.objects.size,
            _ptr(self.cluster_pointers), self.cluster_pointers.size,
            _ptr(self.cluster_idx_to_object), self.cluster_idx_to_object.size,
            _ptr(self.bits), _ptr(self.materials), _ptr(self.luts), self.luts.size // 256,
        )

    def solid_fraction(self) -> float:
        return float(np.unpackbits(self.bits.view(np.uint8)).mean())


def gen_params(kind: int, seed: int, n_objects: int = 0, max_extent: int = 0, world: float = 0.0,
               n_luts: int = 1) -> _GenParams:
    return _GenParams(kind, seed, n_objects, max_extent, world, n_luts)


def generate_scene(params: _GenParams) -> Scene:
    lib = load_library()
    sizes = _GenSizes()
    rc = lib.tgx_gen_scene_sizes(ctypes.byref(params), ctypes.byref(sizes))
    if rc:
        raise TgxError(rc, "tgx_gen_scene_sizes")
    s = Scene(
        objects=np.zeros(sizes.n_objects, OBJECT_DTYPE),
        cluster_pointers=np.zeros(sizes.n_cluster_pointers, np.uint32),
        cluster_idx_to_object=np.zeros(sizes.n_clusters, np.uint32),
        bits=np.zeros(16 * sizes.n_clusters, np.uint32),
        materials=np.zeros(512 * sizes.n_clusters, np.uint8),
        luts=np.zeros(256 * sizes.n_luts, LUT_DTYPE),
    )
    rc = lib.tgx_gen_scene_fill(ctypes.byref(params), _ptr(s.objects), _ptr(s.cluster_pointers),
                                _ptr(s.cluster_idx_to_object), _ptr(s.bits), _ptr(s.materials), _ptr(s.luts))
    if rc:
        raise TgxError(rc, "tgx_gen_scene_fill")
    return s


def save_scene(path, scene: "Scene") -> None:
    """Writes a scene file (tgx_scene_save)."""
    rc = load_library().tgx_scene_save(str(path).encode(), ctypes.byref(scene.desc()))
    if rc:
        raise TgxError(rc, f"tgx_scene_save({path})")


def load_scene(path) -> "Scene":
    """Reads a scene file (tgx_scene_file_counts + tgx_scene_load)."""
    lib = load_library()
    counts = np.zeros(1, SCENE_COUNTS_DTYPE)
    rc = lib.tgx_scene_file_counts(str(path).encode(), _ptr(counts))
    if rc:
        raise TgxError(rc, f"tgx_scene_file_counts({path})")
    c = counts[0]
    s = Scene(objects=np.zeros(int(c["n_objects"]), OBJECT_DTYPE),
              cluster_pointers=np.zeros(int(c["n_cluster_pointers"]), np.uint32),
              cluster_idx_to_object=np.zeros(int(c["n_clusters"]), np.uint32),
              bits=np.zeros(16 * int(c["n_clusters"]), np.uint32),
              materials=np.zeros(512 * int(c["n_clusters"]), np.uint8),
              luts=np.zeros(256 * int(c["n_luts"]), LUT_DTYPE))
    rc = lib.tgx_scene_load(str(path).encode(), _ptr(counts), _ptr(s.objects), _ptr(s.cluster_pointers),
                            _ptr(s.cluster_idx_to_object), _ptr(s.bits), _ptr(s.materials), _ptr(s.luts))
    if rc:
        raise TgxError(rc, f"tgx_scene_load({path})")
    return s


def generate_camera(params: _GenParams, index: int, w: int, h: int) -> np.ndarray:
    cam = np.zeros(1, CAMERA_DTYPE)
    rc = load_library().tgx_gen_camera(ctypes.byref(params), index, w, h, _ptr(cam))
    if rc:
        raise TgxError(rc, "tgx_gen_camera")
    return cam


def make_camera(position, pitch: float, yaw: float, roll: float, fov_y: float, aspect: float,
                near: float = 0.1, far: float = 1000.0) -> np.ndarray:
    cam = np.zeros(1, CAMERA_DTYPE)
    pos = np.ascontiguousarray(position, dtype=np.float32)
    rc = load_library().tgx_camera_make(_ptr(pos), pitch, yaw, roll, fov_y, aspect, near, far, _ptr(cam))
    if rc:
        raise TgxError(rc, "tgx_camera_make")
    return cam


# Workloads of the benchmark and the parity tests.
def single_object_params(seed: int = 1) -> _GenParams:
    """One 4x4x4-cluster object, Bernoulli(0.3) solid bits, one LUT."""
    return gen_params(GEN_SINGLE_BERNOULLI, seed, n_luts=1)


def noise_objects_params(seed: int = 1, n_objects: int = 512, max_extent: int = 8, world: float = 512.0,
                         n_luts: int = 1) -> _GenParams:
    """n_objects noise-filled objects of 1..max_extent clusters per axis."""
    return gen_params(GEN_NOISE_OBJECTS, seed, n_objects, max_extent, world, n_luts)


# --------------------------------------------------------------------------
# GPU context
# --------------------------------------------------------------------------

def decode(word: int) -> tuple[bool, float, int, int]:
    """tgx_visibility_decode: (hit, depth, cluster pointer, voxel idx)."""
    d, c, v = ctypes.c_float(), ctypes.c_uint32(), ctypes.c_uint32()
    hit = load_library().tgx_visibility_decode(int(word), ctypes.byref(d), ctypes.byref(c), ctypes.byref(v))
    return bool(hit), d.value, c.value, v.value


class Context:
    """Owns one tgx_context (one HIP device, one stream)."""

    def __init__(self, device: int = 0):
        self.lib = load_library()
        h = ctypes.c_void_p()
        rc = self.lib.tgx_create(device, ctypes.byref(h))
        if rc:
            raise TgxError(rc, f"tgx_create(device={device})")
        self.h = h

    def close(self) -> None:
        if self.h:
            self.lib.tgx_destroy(self.h)
            self.h = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _check(self, rc: int, what: str) -> None:
        if rc:
            msg = self.lib.tgx_last_error(self.h)
            raise TgxError(rc, f"{what}: {msg.decode() if msg else ''}")

    def upload(self, scene: Scene) -> None:
        self._scene_ref = scene  # keep arrays alive while uploading
        desc = scene.desc()
        self._check(self.lib.tgx_scene_upload(self.h, ctypes.byref(desc)), "tgx_scene_upload")

    def generate_scene_device(self, params: _GenParams) -> None:
        """Builds the scene of `params` in device memory (tgx_gen_scene_device)."""
        self._scene_ref = None
        self._check(self.lib.tgx_gen_scene_device(self.h, ctypes.byref(params)), "tgx_gen_scene_device")

    def scene_export(self, params: _GenParams) -> Scene:
        """The context's scene arrays, sized by the generator parameters."""
        sizes = _GenSizes()
        rc = self.lib.tgx_gen_scene_sizes(ctypes.byref(params), ctypes.byref(sizes))
        if rc:
            raise TgxError(rc, "tgx_gen_scene_sizes")
        s = Scene(objects=np.zeros(sizes.n_objects, OBJECT_DTYPE),
                  cluster_pointers=np.zeros(sizes.n_cluster_pointers, np.uint32),
                  cluster_idx_to_object=np.zeros(sizes.n_clusters, np.uint32),
                  bits=np.zeros(16 * sizes.n_clusters, np.uint32),
                  materials=np.zeros(512 * sizes.n_clusters, np.uint8),
                  luts=np.zeros(256 * sizes.n_luts, LUT_DTYPE))
        self._check(self.lib.tgx_scene_export(self.h, _ptr(s.objects), _ptr(s.cluster_pointers),
                                              _ptr(s.cluster_idx_to_object), _ptr(s.bits), _ptr(s.materials),
                                              _ptr(s.luts)), "tgx_scene_export")
        return s

    def visibility(self, cams: np.ndarray, w: int, h: int) -> np.ndarray:
        cams = np.ascontiguousarray(cams, dtype=CAMERA_DTYPE)
        out = np.empty((cams.size, h, w), np.uint64)
        self._check(self.lib.tgx_visibility(self.h, _ptr(cams), cams.size, w, h, _ptr(out)), "tgx_visibility")
        return out

    def visibility_device(self, cams: np.ndarray, w: int, h: int, out_dev: int) -> None:
        cams = np.ascontiguousarray(cams, dtype=CAMERA_DTYPE)
        self._check(self.lib.tgx_visibility_device(self.h, _ptr(cams), cams.size, w, h, out_dev),
                    "tgx_visibility_device")

    def visibility_device_noalloc(self, cams: np.ndarray, w: int, h: int, out_dev: int) -> None:
        """tgx_visibility_device_noalloc: the same words, work buffers in the
        library's fixed per-GPU scratch array (no device allocation)."""
        cams = np.ascontiguousarray(cams, dtype=CAMERA_DTYPE)
        self._check(self.lib.tgx_visibility_device_noalloc(self.h, _ptr(cams), cams.size, w, h, out_dev),
                    "tgx_visibility_device_noalloc")

    def last_stats(self) -> VisibilityStats:
        st = VisibilityStats()
        self._check(self.lib.tgx_visibility_last_stats(self.h, ctypes.byref(st)), "tgx_visibility_last_stats")
        return st

    def set_list_capacity(self, entries: int) -> None:
        self._check(self.lib.tgx_visibility_set_list_capacity(self.h, entries), "tgx_visibility_set_list_capacity")

    # -- sparse voxel octree ------------------------------------------------

    def svo_build(self, mn, depth: int, flags: int = 0) -> None:
        m = np.ascontiguousarray(mn, np.float32)
        self._check(self.lib.tgx_svo_build(self.h, _ptr(m), depth, flags), "tgx_svo_build")

    def lut_set(self, lut_idx: int, entry: int, albedo, roughness: float) -> None:
        """tgx_lut_set: one LUT entry (tg_raytracer_color_lut_set)."""
        e = np.zeros(1, LUT_DTYPE)
        e["albedo"] = albedo
        e["roughness"] = roughness
        self._check(self.lib.tgx_lut_set(self.h, lut_idx, entry, _ptr(e)), "tgx_lut_set")

    def svo_update(self) -> tuple[int, int]:
        """tgx_svo_update: rebuild after scene edits; (reused, voxelized) leaves."""
        r = np.zeros(2, np.uint32)
        self._check(self.lib.tgx_svo_update(self.h, _ptr(r[0:]), _ptr(r[1:])), "tgx_svo_update")
        return int(r[0]), int(r[1])

    def svo_info(self):
        i = np.zeros(1, SVO_INFO_DTYPE)
        self._check(self.lib.tgx_svo_get_info(self.h, _ptr(i)), "tgx_svo_get_info")
        return i[0]

    def svo_export(self) -> dict:
        info = self.svo_info()
        nl = int(info["n_leaves"])
        out = {
            "info": info,
            "nodes": np.zeros(int(info["n_nodes"]), SVO_NODE_DTYPE),
            "bits": np.zeros((nl, 1024), np.uint32),
            "materials": np.zeros((nl, 32768), np.uint16),
            "leaf_n": np.zeros(nl, np.uint32),
            "list_offsets": np.zeros(nl + 1, np.uint32),
            "list": np.zeros(max(1, int(info["n_list_entries"])), np.uint32),
        }
        self._check(self.lib.tgx_svo_export(self.h, _ptr(out["nodes"]), _ptr(out["bits"]), _ptr(out["materials"]),
                                            _ptr(out["leaf_n"]), _ptr(out["list_offsets"]), _ptr(out["list"])),
                    "tgx_svo_export")
        return out

    def svo_save(self, path) -> None:
        """tgx_svo_save: the built SVO to a file."""
        self._check(self.lib.tgx_svo_save(self.h, str(path).encode()), "tgx_svo_save")

    def svo_load(self, path) -> None:
        """tgx_svo_load: an SVO file saved from the uploaded scene, in place of a build."""
        self._check(self.lib.tgx_svo_load(self.h, str(path).encode()), "tgx_svo_load")

    def svo_trace(self, origins: np.ndarray, dirs: np.ndarray, max_distance: float = 3.0e38) -> np.ndarray:
        o = np.ascontiguousarray(origins, np.float32).reshape(-1, 3)
        d = np.ascontiguousarray(dirs, np.float32).reshape(-1, 3)
        out = np.zeros(o.shape[0], SVO_HIT_DTYPE)
        self._check(self.lib.tgx_svo_trace(self.h, _ptr(o), _ptr(d), o.shape[0], max_distance, _ptr(out)),
                    "tgx_svo_trace")
        return out

    # -- one-bounce GI -------------------------------------------------------

    def gi_device(self, cam: np.ndarray, w: int, h: int, vis_dev: int, params: np.ndarray, radiance_dev: int,
                  irradiance_dev: int | None = None) -> None:
        cam = np.ascontiguousarray(cam, dtype=CAMERA_DTYPE)
        params = np.ascontiguousarray(params, dtype=GI_PARAMS_DTYPE)
        self._check(self.lib.tgx_gi_device(self.h, _ptr(cam), w, h, vis_dev, _ptr(params), radiance_dev,
                                           irradiance_dev), "tgx_gi_device")

    def frame_device(self, cam: np.ndarray, w: int, h: int, params: np.ndarray, vis_dev: int, radiance_dev: int,
                     irradiance_dev: int | None = None) -> None:
        cam = np.ascontiguousarray(cam, dtype=CAMERA_DTYPE)
        params = np.ascontiguousarray(params, dtype=GI_PARAMS_DTYPE)
        self._check(self.lib.tgx_frame_device(self.h, _ptr(cam), w, h, _ptr(params), vis_dev, radiance_dev,
                                              irradiance_dev), "tgx_frame_device")

    def gi_stats(self):
        st = np.zeros(1, GI_STATS_DTYPE)
        self._check(self.lib.tgx_gi_last_stats(self.h, _ptr(st)), "tgx_gi_last_stats")
        return st[0]

    # -- neural irradiance cache ----------------------------------------------

    def nrc_create(self, config: np.ndarray) -> None:
        config = np.ascontiguousarray(config, dtype=NRC_CONFIG_DTYPE)
        self._check(self.lib.tgx_nrc_create(self.h, _ptr(config)), "tgx_nrc_create")

    def nrc_params(self) -> np.ndarray:
        p = np.zeros(NRC_N_PARAMS, np.float32)
        self._check(self.lib.tgx_nrc_get_params(self.h, _ptr(p)), "tgx_nrc_get_params")
        return p

    def nrc_inference_params(self) -> np.ndarray:
        """tgx_nrc_get_inference_params: the weights inference uses (the moving average with ema_decay)."""
        p = np.zeros(NRC_N_PARAMS, np.float32)
        self._check(self.lib.tgx_nrc_get_inference_params(self.h, _ptr(p)), "tgx_nrc_get_inference_params")
        return p

    def nrc_set_params(self, params: np.ndarray) -> None:
        p = np.ascontiguousarray(params, np.float32)
        assert p.size == NRC_N_PARAMS
        self._check(self.lib.tgx_nrc_set_params(self.h, _ptr(p)), "tgx_nrc_set_params")

    @staticmethod
    def _nrc_samples(p, n, albedo) -> np.ndarray:
        return np.ascontiguousarray(np.concatenate([np.asarray(p, np.float32).reshape(-1, 3),
                                                    np.asarray(n, np.float32).reshape(-1, 3),
                                                    np.asarray(albedo, np.float32).reshape(-1, 3)], 1))

    def nrc_infer(self, p, n, albedo) -> np.ndarray:
        """Cache irradiance [N, 3] (clamped at 0) for positions, normals, albedos [N, 3]."""
        x = self._nrc_samples(p, n, albedo)
        out = np.zeros((x.shape[0], 3), np.float32)
        self._check(self.lib.tgx_nrc_infer(self.h, _ptr(x), x.shape[0], _ptr(out)), "tgx_nrc_infer")
        return out

    def nrc_train_batch(self, p, n, albedo, targets, want_grad: bool = False):
        """One Adam step; returns (loss, gradient before the step or None)."""
        x = self._nrc_samples(p, n, albedo)
        t = np.ascontiguousarray(targets, np.float32).reshape(-1, 3)
        loss = ctypes.c_float()
        g = np.zeros(NRC_N_PARAMS, np.float32) if want_grad else None
        self._check(self.lib.tgx_nrc_train_batch(self.h, _ptr(x), _ptr(t), x.shape[0], ctypes.byref(loss), _ptr(g)),
                    "tgx_nrc_train_batch")
        return float(loss.value), g

    def nrc_frame_device(self, cam: np.ndarray, w: int, h: int, params: np.ndarray, vis_dev: int, radiance_dev: int,
                         irradiance_dev: int | None = None) -> None:
        cam = np.ascontiguousarray(cam, dtype=CAMERA_DTYPE)
        params = np.ascontiguousarray(params, dtype=GI_PARAMS_DTYPE)
        self._check(self.lib.tgx_nrc_frame_device(self.h, _ptr(cam), w, h, _ptr(params), vis_dev, radiance_dev,
                                                  irradiance_dev), "tgx_nrc_frame_device")

    def nrc_save(self, path) -> None:
        """Training state of the cache to a file (tgx_nrc_save)."""
        self._check(self.lib.tgx_nrc_save(self.h, str(path).encode()), "tgx_nrc_save")

    def nrc_load(self, path) -> None:
        """Re-creates the cache from a training-state file (tgx_nrc_load)."""
        self._check(self.lib.tgx_nrc_load(self.h, str(path).encode()), "tgx_nrc_load")

    def nrc_stats(self):
        st = np.zeros(1, NRC_STATS_DTYPE)
        self._check(self.lib.tgx_nrc_last_stats(self.h, _ptr(st)), "tgx_nrc_last_stats")
        return st[0]

    def nrc_frame(self, cam: np.ndarray, w: int, h: int, params: np.ndarray):
        """One cache frame: (vis[h, w], radiance[h, w, 3], irradiance[h, w, 3])."""
        vis_d, rad_d, irr_d = self.alloc(8 * w * h), self.alloc(12 * w * h), self.alloc(12 * w * h)
        try:
            self.nrc_frame_device(cam, w, h, params, vis_d, rad_d, irr_d)
            vis = np.empty((h, w), np.uint64)
            rad = np.empty((h, w, 3), np.float32)
            irr = np.empty((h, w, 3), np.float32)
            self.d2h(vis, vis_d)
            self.d2h(rad, rad_d)
            self.d2h(irr, irr_d)
        finally:
            for q in (vis_d, rad_d, irr_d):
                self.free(q)
        return vis, rad, irr

    def frame(self, cam: np.ndarray, w: int, h: int, params: np.ndarray):
        """Visibility + GI of one camera: (vis[h, w], radiance[h, w, 3], irradiance[h, w, 3])."""
        vis_d, rad_d, irr_d = self.alloc(8 * w * h), self.alloc(12 * w * h), self.alloc(12 * w * h)
        try:
            self.frame_device(cam, w, h, params, vis_d, rad_d, irr_d)
            vis = np.empty((h, w), np.uint64)
            rad = np.empty((h, w, 3), np.float32)
            irr = np.empty((h, w, 3), np.float32)
            self.d2h(vis, vis_d)
            self.d2h(rad, rad_d)
            self.d2h(irr, irr_d)
        finally:
            for p in (vis_d, rad_d, irr_d):
                self.free(p)
        return vis, rad, irr

    # -- dynamic scenes ---------------------------------------------------------

    def scene_reserve(self, object_capacity: int, cluster_capacity: int, luts: np.ndarray) -> None:
        luts = np.ascontiguousarray(luts, dtype=LUT_DTYPE)
        assert luts.size % 256 == 0
        self._check(self.lib.tgx_scene_reserve(self.h, object_capacity, cluster_capacity, _ptr(luts), luts.size // 256),
                    "tgx_scene_reserve")

    def object_create(self, n_clusters_per_dim, translation, rotation, lut_idx: int = 0, bits=None,
                      materials=None) -> int:
        n = np.ascontiguousarray(n_clusters_per_dim, np.uint32)
        t = np.ascontiguousarray(translation, np.float32)
        r = np.ascontiguousarray(rotation, np.float32)
        b = None if bits is None else np.ascontiguousarray(bits, np.uint32)
        m = None if materials is None else np.ascontiguousarray(materials, np.uint8)
        k = int(np.prod(n.astype(np.uint64)))
        assert b is None or b.size == 16 * k
        assert m is None or m.size == 512 * k
        out = ctypes.c_uint32()
        self._check(self.lib.tgx_object_create(self.h, _ptr(n), _ptr(t), _ptr(r), lut_idx, _ptr(b), _ptr(m),
                                               ctypes.byref(out)), "tgx_object_create")
        return int(out.value)

    def object_destroy(self, object_idx: int) -> None:
        self._check(self.lib.tgx_object_destroy(self.h, object_idx), "tgx_object_destroy")

    def scene_counts(self):
        c = np.zeros(1, SCENE_COUNTS_DTYPE)
        self._check(self.lib.tgx_scene_get_counts(self.h, _ptr(c)), "tgx_scene_get_counts")
        return c[0]

    def scene_arrays(self) -> Scene:
        """The context's scene arrays (tgx_scene_export, sized by tgx_scene_get_counts)."""
        c = self.scene_counts()
        s = Scene(objects=np.zeros(int(c["n_objects"]), OBJECT_DTYPE),
                  cluster_pointers=np.zeros(int(c["n_cluster_pointers"]), np.uint32),
                  cluster_idx_to_object=np.zeros(int(c["n_clusters"]), np.uint32),
                  bits=np.zeros(16 * int(c["n_clusters"]), np.uint32),
                  materials=np.zeros(512 * int(c["n_clusters"]), np.uint8),
                  luts=np.zeros(256 * int(c["n_luts"]), LUT_DTYPE))
        self._check(self.lib.tgx_scene_export(self.h, _ptr(s.objects), _ptr(s.cluster_pointers),
                                              _ptr(s.cluster_idx_to_object), _ptr(s.bits), _ptr(s.materials),
                                              _ptr(s.luts)), "tgx_scene_export")
        return s

    def svo_device_bytes(self) -> tuple[int, int]:
        """(device bytes of the SVO's data, solid voxels) (tgx_svo_device_bytes)."""
        b, n = ctypes.c_uint64(), ctypes.c_uint64()
        self._check(self.lib.tgx_svo_device_bytes(self.h, ctypes.byref(b), ctypes.byref(n)), "tgx_svo_device_bytes")
        return int(b.value), int(n.value)

    def svo_visibility_device(self, cam: np.ndarray, w: int, h: int, vis_dev: int) -> None:
        cam = np.ascontiguousarray(cam, dtype=CAMERA_DTYPE)
        self._check(self.lib.tgx_svo_visibility_device(self.h, _ptr(cam), w, h, vis_dev), "tgx_svo_visibility_device")

    def svo_visibility(self, cam: np.ndarray, w: int, h: int) -> np.ndarray:
        """SVO primary visibility words [h, w] (tgx_svo_visibility_device)."""
        d = self.alloc(8 * w * h)
        try:
            self.svo_visibility_device(cam, w, h, d)
            out = np.empty((h, w), np.uint64)
            self.d2h(out, d)
        finally:
            self.free(d)
        return out

    # -- pick readback ----------------------------------------------------------

    def pick(self, vis_dev: int, w: int, h: int, xy, cam: np.ndarray | None = None) -> np.ndarray:
        """Decodes pixels (x, y) of a device visibility buffer: PICK_DTYPE [n]."""
        xy = np.ascontiguousarray(np.asarray(xy, np.uint32).reshape(-1, 2))
        out = np.zeros(xy.shape[0], PICK_DTYPE)
        cam_a = np.ascontiguousarray(cam, dtype=CAMERA_DTYPE) if cam is not None else None
        self._check(self.lib.tgx_pick(self.h, _ptr(cam_a), vis_dev, w, h, _ptr(xy), xy.shape[0], _ptr(out)), "tgx_pick")
        return out

    # -- shading pass ---------------------------------------------------------

    def shade_device(self, cam: np.ndarray, w: int, h: int, vis_dev: int, params: np.ndarray, rgba_dev: int) -> None:
        cam = np.ascontiguousarray(cam, dtype=CAMERA_DTYPE)
        params = np.ascontiguousarray(params, dtype=SHADE_PARAMS_DTYPE)
        self._check(self.lib.tgx_shade_device(self.h, _ptr(cam), w, h, vis_dev, _ptr(params), rgba_dev),
                    "tgx_shade_device")

    def shade(self, cam: np.ndarray, w: int, h: int, vis: np.ndarray, params: np.ndarray) -> np.ndarray:
        """Shades a host visibility buffer [h, w]: RGBA float32 [h, w, 4]."""
        vis = np.ascontiguousarray(vis, np.uint64)
        vis_d, out_d = self.alloc(8 * w * h), self.alloc(16 * w * h)
        try:
            self.h2d(vis_d, vis)
            self.shade_device(cam, w, h, vis_d, params, out_d)
            out = np.empty((h, w, 4), np.float32)
            self.d2h(out, out_d)
        finally:
            self.free(vis_d)
            self.free(out_d)
        return out

    # -- batched frames as captured graphs --------------------------------------

    def frame_graph(self, cams: np.ndarray, w: int, h: int, params: np.ndarray, vis_dev: int, radiance_dev: int,
                    irradiance_dev: int | None = None, flags: int = 0) -> "FrameGraph":
        """Captures a frame of len(cams) cameras (runs it once uncaptured first)."""
        cams = np.ascontiguousarray(np.atleast_1d(cams), dtype=CAMERA_DTYPE)
        params = np.ascontiguousarray(params, dtype=GI_PARAMS_DTYPE)
        g = ctypes.c_void_p()
        self._check(self.lib.tgx_frame_graph_create(self.h, _ptr(cams), cams.shape[0], w, h, _ptr(params), flags,
                                                    vis_dev, radiance_dev, irradiance_dev, ctypes.byref(g)),
                    "tgx_frame_graph_create")
        return FrameGraph(self, g.value)

    def frame_batch(self, cams: np.ndarray, w: int, h: int, params: np.ndarray, vis_dev: int, radiance_dev: int,
                    irradiance_dev: int | None = None, flags: int = 0, stats: bool = False):
        """One uncaptured frame of len(cams) cameras; with stats=True waits and returns its totals."""
        cams = np.ascontiguousarray(np.atleast_1d(cams), dtype=CAMERA_DTYPE)
        params = np.ascontiguousarray(params, dtype=GI_PARAMS_DTYPE)
        st = np.zeros(1, FRAME_STATS_DTYPE) if stats else None
        self._check(self.lib.tgx_frame_batch(self.h, _ptr(cams), cams.shape[0], w, h, _ptr(params), flags, vis_dev,
                                             radiance_dev, irradiance_dev, _ptr(st)), "tgx_frame_batch")
        return st[0] if stats else None

    def autotune(self, cams: np.ndarray, w: int, h: int, params: np.ndarray, vis_dev: int, radiance_dev: int,
                 irradiance_dev: int | None = None, flags: int = 0, frames: int = 2) -> np.ndarray:
        """Times the frame over the tuning candidates; keeps and returns the fastest."""
        cams = np.ascontiguousarray(np.atleast_1d(cams), dtype=CAMERA_DTYPE)
        params = np.ascontiguousarray(params, dtype=GI_PARAMS_DTYPE)
        best = np.zeros(1, TUNING_DTYPE)
        self._check(self.lib.tgx_autotune(self.h, _ptr(cams), cams.shape[0], w, h, _ptr(params), flags, vis_dev,
                                          radiance_dev, irradiance_dev, frames, _ptr(best)), "tgx_autotune")
        return best

    def alloc(self, nbytes: int) -> int:
        p = ctypes.c_void_p()
        self._check(self.lib.tgx_device_alloc(self.h, nbytes, ctypes.byref(p)), "tgx_device_alloc")
        return p.value

    def tuning(self) -> np.ndarray:
        t = np.zeros(1, TUNING_DTYPE)
        self._check(self.lib.tgx_tuning_get(self.h, _ptr(t)), "tgx_tuning_get")
        return t

    def set_tuning(self, **kw) -> None:
        t = self.tuning()
        for k, v in kw.items():
            t[k] = v
        self._check(self.lib.tgx_tuning_set(self.h, _ptr(t)), "tgx_tuning_set")

    def d2d(self, dst: int, src: int, nbytes: int) -> None:
        self._check(self.lib.tgx_memcpy_d2d(self.h, dst, src, nbytes), "tgx_memcpy_d2d")

    def free(self, ptr: int) -> None:
        self._check(self.lib.tgx_device_free(self.h, ptr), "tgx_device_free")

    def d2h(self, dst: np.ndarray, src_dev: int) -> None:
        self._check(self.lib.tgx_memcpy_d2h(self.h, _ptr(dst), src_dev, dst.nbytes), "tgx_memcpy_d2h")

    def h2d(self, dst_dev: int, src: np.ndarray) -> None:
        self._check(self.lib.tgx_memcpy_h2d(self.h, dst_dev, _ptr(src), src.nbytes), "tgx_memcpy_h2d")

    def synchronize(self) -> None:
        self._check(self.lib.tgx_synchronize(self.h), "tgx_synchronize")

    def profile(self, enable: bool = True) -> None:
        self._check(self.lib.tgx_profile_enable(self.h, int(enable)), "tgx_profile_enable")

    def profile_reset(self) -> None:
        self._check(self.lib.tgx_profile_reset(self.h), "tgx_profile_reset")

    def profile_stats(self) -> dict[str, tuple[int, float]]:
        n = ctypes.c_uint32()
        buf = (KernelStat * 64)()
        self._check(self.lib.tgx_profile_get(self.h, buf, 64, ctypes.byref(n)), "tgx_profile_get")
        return {buf[i].name.decode(): (int(buf[i].launches), float(buf[i].total_ms)) for i in range(min(n.value, 64))}


class FrameGraph:
    """A captured multi-camera frame (tgx_frame_graph_*)."""

    def __init__(self, ctx: Context, handle: int):
        self.ctx, self.handle = ctx, handle

    def launch(self) -> None:
        self.ctx._check(self.ctx.lib.tgx_frame_graph_launch(self.ctx.h, self.handle), "tgx_frame_graph_launch")

    def stats(self):
        st = np.zeros(1, FRAME_STATS_DTYPE)
        self.ctx._check(self.ctx.lib.tgx_frame_graph_stats(self.ctx.h, self.handle, _ptr(st)), "tgx_frame_graph_stats")
        return st[0]

    def destroy(self) -> None:
        if self.handle and self.ctx.h:
            self.ctx._check(self.ctx.lib.tgx_frame_graph_destroy(self.ctx.h, self.handle), "tgx_frame_graph_destroy")
            self.handle = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.destroy()
