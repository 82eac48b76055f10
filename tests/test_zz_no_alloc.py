"""tgx_visibility_device_noalloc: the allocation-free visibility entry point
gives the same words as tgx_visibility_device, through the C ABI.

Two calls in a row, the second with a smaller input (fewer cameras, lower
resolution, a smaller scene), so that the shared scratch array holds the
first call's contents when the second starts; a call with more cameras than
one batch holds; and a call too large for the scratch array is refused with
TGX_ERR_CAPACITY before anything is launched."""
from __future__ import annotations

import os
import sys

import numpy as np
import pytest

import oracle as O
import tg_amd as T

pytestmark = pytest.mark.gpu


def say(msg: str) -> None:
    print(f"[gpu] {msg}", file=sys.stderr, flush=True)


def both(ctx, cams, w, h):
    """(tgx_visibility_device, tgx_visibility_device_noalloc) words."""
    n = cams.size
    a, b = ctx.alloc(8 * n * w * h), ctx.alloc(8 * n * w * h)
    try:
        ctx.visibility_device(cams, w, h, a)
        ctx.visibility_device_noalloc(cams, w, h, b)
        va = np.empty((n, h, w), np.uint64)
        vb = np.empty((n, h, w), np.uint64)
        ctx.d2h(va, a)
        ctx.d2h(vb, b)
    finally:
        ctx.free(a)
        ctx.free(b)
    return va, vb


def test_noalloc_visibility_equals_allocating_entry_point():
    ctx = T.Context(int(os.environ.get("TGX_DEVICE", "0")))
    try:
        p = T.noise_objects_params(seed=1)
        sc = T.generate_scene(p)
        ctx.upload(sc)
        say("no-alloc visibility: 512 noise objects, 3 cameras at 1920x1080")
        cams = np.concatenate([T.generate_camera(p, i, 1920, 1080) for i in range(3)])
        va, vb = both(ctx, cams, 1920, 1080)
        assert np.array_equal(va, vb)
        assert (va != T.VISIBILITY_EMPTY).mean() > 0.3
        ref, _ = O.visibility_culled(O.OracleScene(sc), cams[1:2], 1920, 1080, row_step=16)
        assert np.array_equal(ref[::16], vb[1, ::16])

        say("no-alloc visibility: second call, one camera at 480x270")
        cam = T.generate_camera(p, 5, 480, 270)
        va, vb = both(ctx, cam, 480, 270)
        assert np.array_equal(va, vb) and (vb != T.VISIBILITY_EMPTY).any()

        say("no-alloc visibility: 18 cameras at 320x180 (two batches)")
        cams = np.concatenate([T.generate_camera(p, i, 320, 180) for i in range(18)])
        va, vb = both(ctx, cams, 320, 180)
        assert np.array_equal(va, vb)

        say("no-alloc visibility: smaller scene (one object), 256x256")
        p1 = T.single_object_params(1)
        s1 = T.generate_scene(p1)
        ctx.upload(s1)
        cam = T.generate_camera(p1, 0, 256, 256)
        va, vb = both(ctx, cam, 256, 256)
        assert np.array_equal(va, vb)
        assert np.array_equal(vb[0], O.visibility_brute(O.OracleScene(s1), cam, 256, 256))

        say("no-alloc visibility: a camera too large for the scratch array is refused")
        w, h = 65535, 512
        big = T.generate_camera(p1, 0, w, h)
        d = ctx.alloc(8 * w * h)
        try:
            with pytest.raises(T.TgxError) as e:
                ctx.visibility_device_noalloc(big, w, h, d)
            assert e.value.code == 2
        finally:
            ctx.free(d)
    finally:
        ctx.close()
