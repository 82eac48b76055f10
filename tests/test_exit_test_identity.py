"""The SVO trace kernel decides "exit distance > FLT_EPSILON" (the pop test
of tg_svo_traverse, tg_sparse_voxel_octree.c:724-733) without dividing:
fl(a / d) > 2^-23  <=>  a > d * 2^-23 (d > 0), a < d * 2^-23 (d < 0), for
|d| >= 2^-100. This checks the identity against IEEE float32 division on
values packed around the rounding boundary and on random values."""
import numpy as np

EPS = np.float32(2.0 ** -23)


def div_free(a, d):
    t = d * np.float32(2.0 ** -23)
    return np.where(d > 0, a > t, a < t)


def by_division(a, d):
    with np.errstate(over="ignore", divide="ignore", invalid="ignore"):
        return (a / d) > EPS


def test_identity_near_boundary():
    rng = np.random.default_rng(7)
    n = 400_000
    # |d| spans many binades, including tiny components down to 2^-100
    e = rng.uniform(-100, 1, n)
    d = (rng.uniform(1, 2, n) * 2.0 ** e).astype(np.float32)
    d = np.where(rng.random(n) < 0.5, -d, d).astype(np.float32)
    # a placed at the boundary a ~ d * 2^-23 * (1 + 2^-24), then moved by -3..3 ulps
    a0 = (d.astype(np.float64) * 2.0 ** -23 * (1 + 2.0 ** -24)).astype(np.float32)
    for k in range(-3, 4):
        a = a0.copy()
        for _ in range(abs(k)):
            a = np.nextafter(a, np.float32(np.inf) if k > 0 else np.float32(-np.inf)).astype(np.float32)
        assert np.array_equal(div_free(a, d), by_division(a, d)), k


def test_identity_exact_midpoints():
    # a / d exactly equal to the midpoint 2^-23 (1 + 2^-24): ties go to the even 2^-23, so "not greater"
    m = np.arange(1, 1 << 12, dtype=np.int64) * 2 + 1  # odd 13-bit mantissas
    d = (m.astype(np.float64) * 2.0 ** -10).astype(np.float32)
    a64 = d.astype(np.float64) * 2.0 ** -23 * (1 + 2.0 ** -24)
    exact = a64.astype(np.float32).astype(np.float64) == a64
    a = a64.astype(np.float32)[exact]
    dd = d[exact]
    assert np.array_equal(div_free(a, dd), by_division(a, dd))


def test_identity_random():
    rng = np.random.default_rng(11)
    n = 1_000_000
    d = rng.standard_normal(n).astype(np.float32)
    a = (rng.standard_normal(n) * 10.0 ** rng.integers(-9, 3, n)).astype(np.float32)
    ok = np.abs(d) >= np.float32(2.0 ** -100)
    assert np.array_equal(div_free(a[ok], d[ok]), by_division(a[ok], d[ok]))
