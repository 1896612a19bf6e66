# SPDX-License-Identifier: Apache-2.0
"""The fused 2-D cube (both passes in one kernel for windows up to 31 samples, the column-filtered
rows held in LDS) bitwise against the oracle's cube: every window from 1 to 35 in one launch (both
fused classes, 1..15 and 17..31, and the two-pass path for 33 and 35, which skips the fused
elements), even windows (the same odd side), idle elements (window 0), fields of 128 to 1000 rows
and 64 to 512 columns (column counts that are not a multiple of 64 leave lanes without a column),
max, min, box, max + min, float32 and float64; fields that do not qualify (fewer than 128 rows or 64
columns, in place) keep the two-pass path."""

from __future__ import annotations

import zlib

import numpy as np
import pytest

import gpu_util as G
import oracle

pytestmark = pytest.mark.gpu
DTYPES = [np.float32, np.float64]
WINDOWS = list(range(1, 36, 2)) + [0]


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
        x[..., ::5] = -0.0
    return x.astype(dtype)


def check(x, op, wins, got):
    for b, w in enumerate(wins):
        if w == 0:
            continue
        if op == "maxmin":
            assert bits_equal(got[0][b], oracle.cube(x[b], "max", w)), (w, "max")
            assert bits_equal(got[1][b], oracle.cube(x[b], "min", w)), (w, "min")
        else:
            assert bits_equal(got[b], oracle.cube(x[b], op, w)), w


@pytest.mark.parametrize("dtype", DTYPES)
@pytest.mark.parametrize("shape", [(128, 64), (256, 512), (300, 448), (200, 100), (1000, 65), (129, 511)])
@pytest.mark.parametrize("op", ["max", "min", "box", "maxmin"])
def test_every_window_is_bitwise_the_oracle(dtype, shape, op):
    print(f"fused cube: {op} {shape} {np.dtype(dtype).name}", flush=True)
    x = field((len(WINDOWS),) + shape, dtype, zlib.crc32(repr((shape, op)).encode()), ties=op != "box")
    check(x, op, WINDOWS, G.cube(x, op, WINDOWS))


@pytest.mark.parametrize("op", ["max", "box", "maxmin"])
def test_one_class_per_launch(op):
    """Launches whose elements all fall in one class (the other class and the two-pass path are
    not launched), and one whose elements all take the two-pass path."""
    shape = (256, 256)
    for wins in ([5, 5, 7, 1], [17, 31, 25, 19], [33, 65, 201, 0], [31, 33]):
        x = field((len(wins),) + shape, np.float32, zlib.crc32(repr((op, wins)).encode()))
        check(x, op, wins, G.cube(x, op, wins))


@pytest.mark.parametrize("op", ["max", "box"])
def test_fields_outside_the_fused_shapes(op):
    """Fewer than 128 rows, fewer than 64 or more than 512 columns, and the in-place cube: the
    two-pass path, the same bits."""
    for shape in [(127, 256), (256, 63), (128, 600)]:
        wins = [3, 15, 31]
        x = field((len(wins),) + shape, np.float32, zlib.crc32(repr((op, shape)).encode()))
        check(x, op, wins, G.cube(x, op, wins))
    wins = [3, 15, 31]
    x = field((3, 256, 256), np.float32, 7)
    check(x, op, wins, G.cube(x, op, wins, inplace=True))
