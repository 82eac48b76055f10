"""Loads the golden fixtures of tests/golden and rebuilds their inputs."""
from __future__ import annotations

import importlib.util
from pathlib import Path

import numpy as np

import tg_amd as T

GOLDEN = Path(__file__).resolve().parent / "golden"

_spec = importlib.util.spec_from_file_location("make_golden", GOLDEN / "make_golden.py")
make = importlib.util.module_from_spec(_spec)
_spec.loader.exec_module(make)


def load(name: str) -> dict:
    with np.load(GOLDEN / name, allow_pickle=False) as f:
        return {k: f[k] for k in f.files}


def camera(fix: dict) -> np.ndarray:
    return fix["camera"].view(T.CAMERA_DTYPE).reshape(1)


def hits(fix: dict) -> np.ndarray:
    return np.ascontiguousarray(fix["hits"]).view(T.SVO_HIT_DTYPE).reshape(-1)
