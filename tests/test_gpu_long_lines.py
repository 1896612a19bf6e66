# SPDX-License-Identifier: Apache-2.0
"""Lines the staged kernels do not hold in LDS (GPU): the cube filter's global-memory pass keeps the
canonical vHGW order, so its max / min / box results are bitwise the oracle's on 1-D lines of up to
2^18 samples (also in place: input and output in the same memory) and on lines of 8192 (float32) /
4096 (float64) samples with box windows up to the whole reflected line (2n + 1); 1-D decompositions of long signals (the reference stops at 8192
samples) against the FP64 oracle: decisions exact, rows within 1e-11 relative L2 in float64. In
float32 the bound is float32's own sensitivity on these signals: at n = 12000 the oracle run in
float32 differs from its float64 run by 2.9e-5 relative L2 in mode 5, and the GPU's float32 rows by
the same (bound 1e-4); at n = 65536 an extremum test of mode 4 sits on a near tie that the GPU's
float32 residual (within 3.2e-6 of the oracle's float32 one) decides the other way, which moves the
scale field near that sample: 9.3e-4 in mode 4, 1.7e-3 in mode 5 (bound 2e-3; the float64 run
agrees to 1e-17). The tape of every mode (the frozen decisions) against the oracle's tape; the
stages of a mode (energy, noise level, quantiles, threshold field, extrema flags, raw window,
components) bitwise on 65536-sample lines."""

from __future__ import annotations

import zlib

import numpy as np
import pytest

import eda
import gpu_util as G
import oracle

pytestmark = pytest.mark.gpu


def bits_equal(a, b):
    a = np.ascontiguousarray(a)
    b = np.ascontiguousarray(b, dtype=a.dtype)
    return np.array_equal(a.view(np.uint8), b.view(np.uint8))


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
@pytest.mark.parametrize("n,wins", [(20000, (3, 65, 4097, 40001)), (70001, (9, 1025, 140003))])
@pytest.mark.parametrize("op", ["max", "min", "box"])
def test_long_line_cube_is_bitwise_the_oracle(dtype, n, wins, op):
    print(f"long lines: cube {op} n={n} {np.dtype(dtype).name}", flush=True)
    rng = np.random.default_rng(zlib.crc32(repr((n, op)).encode()))
    x = rng.standard_normal((len(wins), n)).astype(dtype)
    got = G.cube(x, op, list(wins))
    for b, w in enumerate(wins):
        assert bits_equal(got[b], oracle.cube(x[b], op, w)), (n, w)


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
@pytest.mark.parametrize("op", ["max", "box"])
def test_long_line_cube_in_place(dtype, op):
    # the input is the output (as in every pass after a cube's first, and in the decomposition's
    # window and scale fields): the outputs must come from the input as it was
    n, wins = 12000, (11, 39, 319, 827, 24001)
    print(f"long lines: in-place cube {op} n={n} {np.dtype(dtype).name}", flush=True)
    rng = np.random.default_rng(zlib.crc32(repr((n, op, "in place")).encode()))
    x = np.cumsum(rng.standard_normal((len(wins), n)), axis=1).astype(dtype)
    got = G.cube(x, op, list(wins), inplace=True)
    for b, w in enumerate(wins):
        assert bits_equal(got[b], oracle.cube(x[b], op, w)), w


@pytest.mark.parametrize("dtype,n", [(np.float32, 8192), (np.float64, 4096)])
def test_whole_line_box_windows_at_the_staged_limit(dtype, n):
    rng = np.random.default_rng(n)
    wins = [2 * n - 1, 2 * n + 1, (4 * n) // 3 | 1]
    x = rng.standard_normal((len(wins), n)).astype(dtype)
    got = G.cube(x, "box", wins)
    for b, w in enumerate(wins):
        assert bits_equal(got[b], oracle.cube(x[b], "box", w)), w


def long_signal(n, seed):
    rng = np.random.default_rng(seed)
    t = np.arange(n, dtype=np.float64)
    f0, f1 = rng.uniform(0.001, 0.01), rng.uniform(0.05, 0.2)
    om = rng.uniform(0.005, 0.05)
    x = np.cos(2 * np.pi * (f0 * t + (f1 - f0) * t * t / (2 * n))) + np.cos(om * t + 0.8 * np.sin(om * t))
    return x + (t / n) ** 2 + 0.01 * rng.standard_normal(n)


@pytest.mark.parametrize("dtype,n,tol", [(np.float32, 12000, 1e-4), (np.float32, 65536, 2e-3),
                                         (np.float64, 12000, 1e-11), (np.float64, 65536, 1e-11)])
def test_long_1d_decomposition_against_the_oracle(dtype, n, tol):
    print(f"long lines: decompose n={n} {np.dtype(dtype).name}", flush=True)
    x = long_signal(n, n).astype(dtype)
    y, info = eda.decompose(x, return_info=True)
    yo, io = oracle.decompose(x.astype(np.float64))
    assert int(info.n_imfs) == int(io["n_imfs"][0])
    assert info.windows.tolist() == io["windows"][0].tolist()
    assert info.sifts.tolist() == io["sifts"][0].tolist()
    for k in range(int(info.n_imfs) + 1):
        err = float(np.linalg.norm(y[k] - yo[0, k]) / max(np.linalg.norm(yo[0, k]), 1e-300))
        assert err <= tol, (k, err)


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_long_1d_tape_is_the_oracle_tape(dtype):
    # the frozen decisions of every mode against the oracle's run in the same dtype: windows, rung
    # ladders and their used / segment-point bits, sifts run and accepted; the float64 scale
    # fields and floors to 1e-12. In float32 the first four modes' fields agree to 2e-6 relative;
    # from mode 4 on an extremum test that sits on a near tie (sample 38749: its threshold agrees to
    # 6 digits) flags differently, which moves the scale field there.
    import torch

    n = 65536
    x = long_signal(n, n).astype(dtype)
    print(f"long lines: tape n={n} {np.dtype(dtype).name}", flush=True)
    td = eda.TapedDecomposition(torch.from_numpy(x[None]).to("cuda"))
    ref = oracle.decompose_tape(x)
    assert int(td.info["n_imfs"][0]) == ref["n_imfs"]
    for m, om in enumerate(ref["modes"]):
        t = td.mode_tape(0, m)
        for key in ("w", "W", "rung_w", "used", "pts", "ran", "accepted"):
            assert t[key] == om[key], (m, key)
        if dtype == np.float64:
            assert np.allclose(t["thr"], om["thr"], rtol=1e-12, atol=0) and t["tol"] == om["tol"]
            assert np.max(np.abs(t["L"] - om["L"]) / om["L"]) <= 1e-12, m
            assert np.max(np.abs(t["r"] - om["r"])) <= 1e-12 * np.max(np.abs(om["r"])), m
        elif m < 4:
            assert np.max(np.abs(t["L"].astype(np.float64) - om["L"]) / om["L"]) <= 2e-6, m


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_long_line_primitives_are_the_oracle(dtype):
    # the stages of a mode on 65536-sample lines: the signal and the oracle's residual at its mode 4
    n = 65536
    x = long_signal(n, n).astype(dtype)
    tape = oracle.decompose_tape(x)
    m4 = tape["modes"][4]
    xs = np.stack([x, m4["r"]])
    print(f"long lines: primitives n={n} {np.dtype(dtype).name}", flush=True)
    e, a, nf = G.stats(xs)
    for b in range(2):
        assert (e[b], a[b]) == oracle.energy(xs[b]) and nf[b] == 0
    sig = G.noise(xs, False)
    for b in range(2):
        assert sig[b] == oracle.noise_sigma(xs[b], False)
    for q in (0.5, 0.9):
        got, cnt = G.select(np.abs(xs), None, 1, q)
        for b in range(2):
            assert got[b] == oracle.quantile(np.abs(xs[b]).astype(np.float64), q) and cnt[b] == n
    Wd, Wl = [161, 487], [321, 963]
    tol = [1e-6, m4["tol"]]
    tau, tnf = G.tau(xs, Wd, Wl, list(sig), tol)
    fl, ext = G.flags(xs, tau)
    for b in range(2):
        rt, ok = oracle.tau(xs[b], Wd[b], Wl[b], sig[b], tol[b])
        assert bits_equal(tau[b], rt) and tnf[b] == (0 if ok else 1)
        rf, re = oracle.sig_flags(xs[b], rt)
        assert np.array_equal(fl[b], rf)
        assert ext[b, :1].tolist() == [int(v) for v in re]
    wins = [m4["W"], 3 * m4["W"] | 1]
    om = G.raw_window(fl, wins, dtype)
    for b in range(2):
        assert bits_equal(om[b], oracle.raw_window(fl[b], wins[b], dtype)), b
    mask = om > np.median(om[np.isfinite(om)])
    lab = G.ccl(mask)
    for b in range(2):
        assert np.array_equal(lab[b], oracle.ccl(mask[b]))


def test_components_across_line_pieces():
    # 1-D lines are labelled in pieces of 4096 voxels: runs that cross one or several pieces (and a
    # whole line in one run) take the label of their first voxel
    n = 70000
    m = np.ones((4, n), bool)
    m[0, [5000, 20000, 20001, 61439]] = False  # runs over 1, 4 and 10 pieces
    m[1] = np.random.default_rng(1).random(n) < 0.7
    m[2, ::4096] = False  # a gap at every piece start
    m[2, 4095::4096] = False
    got = G.ccl(m)
    for b in range(4):
        assert np.array_equal(got[b], oracle.ccl(m[b])), b
