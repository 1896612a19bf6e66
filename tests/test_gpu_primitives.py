# SPDX-License-Identifier: Apache-2.0
"""Every HIP stage against the oracle, through the C ABI's primitive entry points, on seeded inputs
(GPU). Integer, order and index results are bit-exact; the cube filter's sums, the reductions, the
selections and the thresholds are bit-exact too, in float32 and float64, because the kernels keep
the oracle's order of IEEE operations; the DCT (a GEMM) and the orientation are within the stated
tolerance."""

from __future__ import annotations

import zlib

import numpy as np
import pytest

import gpu_util as G
import oracle

pytestmark = pytest.mark.gpu
DTYPES = [np.float32, np.float64]


def bits_equal(a, b):
    a = np.ascontiguousarray(a)
    b = np.ascontiguousarray(b, dtype=a.dtype)
    return np.array_equal(a.view(np.uint8), b.view(np.uint8))


@pytest.mark.parametrize("dtype", DTYPES)
@pytest.mark.parametrize("shape", [(1000,), (3000,), (37, 70), (9, 10, 11), (512, 512)])
@pytest.mark.parametrize("op", ["max", "min", "box"])
def test_cube_is_bitwise_the_oracle(dtype, shape, op):
    rng = np.random.default_rng(zlib.crc32(repr((shape, op)).encode()))
    B = 3
    x = rng.standard_normal((B,) + shape).astype(dtype)
    for w in (1, 3, 33, 65, 257, 2 * max(shape) + 5):
        got = G.cube(x, op, [w, w, 0] if w > 3 else [w] * B)
        for b in range(B):
            if w > 3 and b == 2:
                continue
            assert bits_equal(got[b], oracle.cube(x[b], op, w)), (w, b)


@pytest.mark.parametrize("dtype", DTYPES)
def test_cube_maxmin_and_per_element_windows(dtype):
    rng = np.random.default_rng(7)
    x = rng.standard_normal((4, 64, 96)).astype(dtype)
    wins = [3, 17, 129, 41]
    mx, mn = G.cube(x, "maxmin", wins)
    for b, w in enumerate(wins):
        assert bits_equal(mx[b], oracle.cube(x[b], "max", w))
        assert bits_equal(mn[b], oracle.cube(x[b], "min", w))


@pytest.mark.parametrize("dtype", DTYPES)
@pytest.mark.parametrize("shape", [(2048,), (300, 301), (20, 30, 40), (1, 8192)])
def test_energy_and_amax_are_bitwise(dtype, shape):
    x = np.random.default_rng(1).standard_normal((2,) + shape).astype(dtype)
    x[1, ..., 3] = np.inf
    e, a, nf = G.stats(x)
    e0, a0 = oracle.energy(x[0])
    assert e[0] == e0 and a[0] == a0 and nf[0] == 0
    assert nf[1] == np.prod(shape) // shape[-1]


@pytest.mark.parametrize("dtype", DTYPES)
@pytest.mark.parametrize("n", [1, 2, 3, 10, 1001, 100000, 300000])
def test_selection_is_numpy(dtype, n):
    """n = 300000: about 90000 zeros in element 0 match the median's leading digits, more than the
    65536 candidate slots an element keeps after the third radix pass (the later passes then read
    every value); element 1 (normal values) counts from its candidate list."""
    rng = np.random.default_rng(n)
    x = rng.standard_normal((3, n)).astype(dtype)
    x[0, : n // 2] = 0.0
    x[0, : n // 4] *= -1.0  # signed zeros count as one value
    m = rng.random((3, n)) < 0.6
    m[2] = False
    med, cnt = G.select(x, m, 0)
    qua, _ = G.select(x, m, 1, 0.25)
    for b in range(3):
        v = x[b][m[b]].astype(np.float64)
        assert cnt[b] == v.size
        if v.size == 0:
            assert np.isnan(med[b]) and np.isnan(qua[b])
            continue
        assert med[b] == np.median(v) and qua[b] == np.quantile(v, 0.25)


@pytest.mark.parametrize("dtype", DTYPES)
@pytest.mark.parametrize("shape", [(2048,), (64, 80), (12, 14, 16)])
def test_noise_levels_are_bitwise(dtype, shape):
    x = np.random.default_rng(2).standard_normal((2,) + shape).astype(dtype)
    for per in (False, True):
        got = G.noise(x, per)
        for b in range(2):
            assert got[b] == oracle.noise_sigma(x[b], per)


@pytest.mark.parametrize("dtype", DTYPES)
@pytest.mark.parametrize("shape", [(4096,), (128, 128), (16, 24, 32)])
def test_threshold_field_is_bitwise(dtype, shape):
    x = np.random.default_rng(3).standard_normal((2,) + shape).astype(dtype)
    Wd, Wl = [7, 13], [11, 25]
    sig = [oracle.noise_sigma(x[b], False) for b in range(2)]
    tol = [1e-4, 2e-4]
    got, nf = G.tau(x, Wd, Wl, sig, tol)
    for b in range(2):
        ref, ok = oracle.tau(x[b], Wd[b], Wl[b], sig[b], tol[b])
        assert bits_equal(got[b], ref) and nf[b] == (0 if ok else 1)


@pytest.mark.parametrize("dtype", DTYPES)
@pytest.mark.parametrize("shape", [(3000,), (5000,), (70, 90), (64, 512), (20, 21, 22), (3, 40)])
@pytest.mark.parametrize("coarse", [False, True])
def test_extrema_flags_are_exact(dtype, shape, coarse):
    rng = np.random.default_rng(4)
    x = rng.standard_normal((2,) + shape)
    x = (np.round(x * 3) if coarse else x).astype(dtype)  # coarse: plateaus of equal values
    t = np.abs(rng.standard_normal(x.shape)).astype(dtype) * 0.5
    fl, ext = G.flags(x, t)
    for b in range(2):
        rf, re = oracle.sig_flags(x[b], t[b])
        assert np.array_equal(fl[b], rf)
        assert ext[b, : len(shape)].tolist() == [int(v) for v in re]


@pytest.mark.parametrize("dtype", DTYPES)
@pytest.mark.parametrize("shape", [(500,), (60, 70), (10, 12, 14)])
def test_raw_window_is_bitwise(dtype, shape):
    rng = np.random.default_rng(5)
    D = len(shape)
    fl = (rng.random((2,) + shape) < 0.1).astype(np.uint8)
    for a in range(1, D):
        fl |= ((rng.random((2,) + shape) < 0.1).astype(np.uint8) << a)
    fl[1, ...] = 0  # no extrema: +inf
    wins = [9, 21]
    got = G.raw_window(fl, wins, dtype)
    for b in range(2):
        assert bits_equal(got[b], oracle.raw_window(fl[b], wins[b], dtype))


@pytest.mark.parametrize("shape", [(1000,), (64, 80), (16, 17, 18), (512, 512)])
def test_components_are_exact(shape):
    m = np.random.default_rng(6).random((2,) + shape) < 0.55
    got = G.ccl(m)
    for b in range(2):
        assert np.array_equal(got[b], oracle.ccl(m[b].astype(np.uint8)))


@pytest.mark.parametrize("dtype,tol", [(np.float32, 3e-6), (np.float64, 1e-13)])
@pytest.mark.parametrize("shape", [(64, 48), (16, 20, 24), (512, 512), (128, 33), (32, 64, 16), (2048,), (33, 4096)])
def test_dct_matches_the_oracle(dtype, tol, shape):
    """Both transforms: the dense matrices (axes of other lengths) and the fast transform of
    power-of-two axes from 16 samples (an odd number of lines pairs the last with a zero line)."""
    x = np.random.default_rng(7).standard_normal((2,) + shape).astype(dtype)
    for inv, odd in ((0, 0), (1, 0), (1, 1), (1, 2), (1, 3)):
        got = G.dct(x, inv, odd)
        for b in range(2):
            ref = oracle.dct(x[b].astype(np.float64), bool(inv), odd)
            assert np.abs(got[b] - ref).max() <= tol * np.abs(ref).max() * max(1.0, np.sqrt(np.prod(shape)) / 64)


@pytest.mark.parametrize("dtype,tol", [(np.float32, 2e-4), (np.float64, 1e-10)])
@pytest.mark.parametrize("shape", [(64, 64), (24, 28, 32)])
def test_orientation_matches_the_oracle(dtype, tol, shape):
    import workloads

    if len(shape) == 2:
        x = workloads.rings_stripes_blobs(2, 64).astype(dtype)
    else:
        x = np.random.default_rng(8).standard_normal((2,) + shape).astype(dtype)
    wins = [9, 15]
    nv, c = G.orientation(x, wins)
    for b in range(2):
        rn, rc, ok = oracle.orientation(x[b].astype(np.float64), wins[b])
        assert ok
        # the coherence within tol where the orientation is defined; n up to its sign
        assert np.abs(c[b] - rc).max() <= tol * 50
        strong = rc > 0.2
        dots = np.abs(sum(nv[a, b] * rn[a] for a in range(len(shape))))
        assert np.all(dots[strong] >= 1 - tol * 100)
