"""The C ABI: libtgx.so loads and exports every function include/*.h
declares; host-only entry points behave as documented (no GPU needed)."""
from __future__ import annotations

import ctypes
import re
from pathlib import Path

import numpy as np

import tg_amd as T

INCLUDE = Path(__file__).resolve().parent.parent / "include"


def declared_functions() -> set[str]:
    names = set()
    for h in INCLUDE.glob("*.h"):
        text = re.sub(r"/\*.*?\*/", "", h.read_text(), flags=re.S)
        for m in re.finditer(r"\b(tgx_[a-z0-9_]+)\s*\(", text):
            names.add(m.group(1))
    return names


def test_library_exports_every_declared_function():
    lib = ctypes.CDLL(str(T.LIB_PATH))
    names = declared_functions()
    assert len(names) >= 20
    missing = [n for n in sorted(names) if not hasattr(lib, n)]
    assert not missing, missing


def test_bindings_cover_the_header():
    T.load_library()
    bound = set(T.load_library().__dict__) | {n for n in declared_functions() if hasattr(T.load_library(), n)}
    assert declared_functions() <= bound


def test_decode_matches_reference_unpack():
    # tgvk_raytracer.c:1639-1653: depth = (word >> 40) / 16777215; a hit iff depth < 1
    w = (1234567 << 40) | (987654 << 9) | 321
    hit, depth, cptr, vox = T.decode(w)
    assert hit and cptr == 987654 and vox == 321
    assert depth == np.float32(np.float32(1234567) / np.float32(16777215.0))
    hit, depth, cptr, vox = T.decode(0xFFFFFFFFFFFFFFFF)
    assert not hit and depth == 1.0 and cptr == 0xFFFFFFFF and vox == 0xFFFFFFFF


def test_null_arguments_are_rejected_without_a_device():
    lib = T.load_library()
    assert lib.tgx_create(0, None) == 1               # TGX_ERR_INVALID_ARG
    assert lib.tgx_destroy(None) == 1
    assert lib.tgx_scene_upload(None, None) == 1
    assert lib.tgx_visibility(None, None, 0, 0, 0, None) == 1
    assert lib.tgx_last_error(None) == b"null context"


def test_tuning_dtype_matches_the_header_struct():
    """tg_amd's tgx_tuning record has the header struct's fields, in order
    (all uint32_t), so tgx_tuning_get / _set exchange the same bytes."""
    text = re.sub(r"/\*.*?\*/", "", (INCLUDE / "tgx.h").read_text(), flags=re.S)
    body = re.search(r"typedef struct tgx_tuning \{(.*?)\} tgx_tuning;", text, flags=re.S).group(1)
    fields = re.findall(r"uint32_t\s+(\w+)\s*;", body)
    assert len(fields) == len(re.findall(r";", body))
    assert list(T.TUNING_DTYPE.names) == fields
    assert T.TUNING_DTYPE.itemsize == 4 * len(fields)
