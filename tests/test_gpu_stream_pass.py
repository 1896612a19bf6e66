# SPDX-License-Identifier: Apache-2.0
"""The streaming pass of strided axes (windows up to 63 samples, lines of at least 128 samples, one
lane per line, rows read through scalar offsets) bitwise against the oracle's cube: every odd window
from 1 to 63 (the short walk to 31, the long one with two canonical chunks per box segment above);
launches that mix them with longer windows (the staged kernels take those and skip
the streamed elements) and with elements that do not run (window 0); slabs whose line count is not
a multiple of 64; shapes whose contiguous axis runs the other kernels; 1-D signals; 3-D volumes
(axes 0 and 1 stream); signed zeros and ties; float32 and float64; the in-place cube (its first
pass writes a scratch field, so axis 0 still streams)."""

from __future__ import annotations

import zlib

import numpy as np
import pytest

import gpu_util as G
import oracle

pytestmark = pytest.mark.gpu
DTYPES = [np.float32, np.float64]
WINDOWS = list(range(1, 64, 2))


def bits_equal(a, b):
    a = np.ascontiguousarray(a)
    b = np.ascontiguousarray(b, dtype=a.dtype)
    return np.array_equal(a.view(np.uint8), b.view(np.uint8))


def field(shape, dtype, seed, ties=False):
    rng = np.random.default_rng(seed)
    x = rng.standard_normal(shape)
    if ties:
        # quantized values: many equal samples, and signed zeros among them
        x = np.round(x * 2.0) / 2.0
        x[..., ::7] = -0.0
    return x.astype(dtype)


@pytest.mark.parametrize("dtype", DTYPES)
@pytest.mark.parametrize(
    "shape", [(128, 64), (200, 50), (512, 130), (130, 9, 70), (96, 256), (40, 300), (3000,), (6, 20, 160)]
)
@pytest.mark.parametrize("op", ["max", "min", "box", "maxmin"])
def test_every_short_window_is_bitwise_the_oracle(dtype, shape, op):
    wins = WINDOWS
    x = field((len(wins),) + shape, dtype, zlib.crc32(repr((shape, op, "w")).encode()), ties=op != "box")
    got = G.cube(x, op, wins)
    for b, w in enumerate(wins):
        if op == "maxmin":
            assert bits_equal(got[0][b], oracle.cube(x[b], "max", w)), (w, "max")
            assert bits_equal(got[1][b], oracle.cube(x[b], "min", w)), (w, "min")
        else:
            assert bits_equal(got[b], oracle.cube(x[b], op, w)), w


@pytest.mark.parametrize("dtype", DTYPES)
@pytest.mark.parametrize("op", ["max", "box", "maxmin"])
def test_mixed_launch_and_idle_elements(dtype, op):
    """Launches that span the short walk, the long walk and the staged kernels (windows above 63),
    each skipping the others' elements; elements with window 0 are not filtered."""
    shape = (256, 96)
    for wins in ([31, 0, 5, 1], [33, 0, 5, 1], [65, 63, 0, 31, 1], [129, 3]):
        x = field((len(wins),) + shape, dtype, zlib.crc32(repr((op, wins)).encode()))
        got = G.cube(x, op, wins)
        for b, w in enumerate(wins):
            if w == 0:
                continue
            if op == "maxmin":
                assert bits_equal(got[0][b], oracle.cube(x[b], "max", w)), (wins, w)
                assert bits_equal(got[1][b], oracle.cube(x[b], "min", w)), (wins, w)
            else:
                assert bits_equal(got[b], oracle.cube(x[b], op, w)), (wins, w)


@pytest.mark.parametrize("dtype", DTYPES)
@pytest.mark.parametrize("op", ["max", "min", "box"])
def test_in_place_cube(dtype, op):
    wins = [3, 15, 31, 45, 63, 65]
    x = field((6, 512, 512), dtype, 11)
    got = G.cube(x, op, wins, inplace=True)
    for b, w in enumerate(wins):
        assert bits_equal(got[b], oracle.cube(x[b], op, w)), w


def test_full_benchmark_shape_float32():
    """64 images of 512 x 512 (the benchmark's shape): the pieces of a line and the chip-filling
    split; every eighth image checked against the oracle."""
    B = 64
    x = field((B, 512, 512), np.float32, 5)
    wins = [1 + 2 * (b % 16) for b in range(B)]
    for op in ("box", "max"):
        got = G.cube(x, op, wins)
        for b in range(0, B, 8):
            assert bits_equal(got[b], oracle.cube(x[b], op, wins[b])), (op, b)
