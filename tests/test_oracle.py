# SPDX-License-Identifier: Apache-2.0
"""The oracle pinned (CPU): against the reference's golden corpus and known-answer tables, and
each primitive against the closed form that fixes it (SciPy / NumPy routines or brute force), so
that a dropped term, a wrong index or a transposed operand fails a check.

Golden rule (reference tests/python/test_golden.py:13-27,119-137): decisions (n_imfs, windows,
sifts) exact; rows within A * 1000 * 2^-53 * max|x| with A = sensitivity / 1e-12."""

from __future__ import annotations

import json
from pathlib import Path

import numpy as np
import pytest
import scipy.fft
import scipy.ndimage
import scipy.signal

import oracle

GOLDEN = Path(__file__).resolve().parent / "golden"
CASES = sorted(p.stem for p in (GOLDEN / "corpus").glob("*.npz"))


def load_case(name):
    with np.load(GOLDEN / "corpus" / f"{name}.npz", allow_pickle=False) as z:
        return {k: z[k] for k in z.files}


# ---- end to end: the golden corpus (FP64) -------------------------------------------------------
@pytest.mark.parametrize("name", CASES)
def test_oracle_reproduces_the_golden_corpus(name):
    case = load_case(name)
    x, modes = case["x"], case["modes"]
    y, info = oracle.decompose(x)
    assert int(info["n_imfs"][0]) == modes.shape[0] - 1
    assert info["windows"][0].tolist() == case["windows"].tolist()
    assert info["sifts"][0].tolist() == case["sifts"].tolist()
    bound = float(case["sensitivity"]) / 1e-12 * 1000 * 2.0**-53 * float(np.abs(x).max())
    assert float(np.abs(y[0] - modes).max()) <= bound


def test_golden_index_matches_the_files():
    index = json.loads((GOLDEN / "corpus" / "index.json").read_text())
    assert sorted(index["cases"]) == CASES
    for name in CASES:
        case = load_case(name)
        assert list(case["x"].shape) == index["cases"][name]["shape"]


# ---- invariants of the decomposition -------------------------------------------------------------
def test_rows_reconstruct_the_input():
    x = np.load(GOLDEN / "corpus" / "tones_2048.npz")["x"]
    y, _ = oracle.decompose(x)
    S = np.abs(y[0]).sum(axis=0).max()
    assert np.abs(y[0].sum(axis=0) - x).max() <= 64 * 2.0**-53 * S


def test_gain_by_a_power_of_two_is_bitwise_equivariant():
    x = np.load(GOLDEN / "corpus" / "amfm_512.npz")["x"]
    y1, i1 = oracle.decompose(x)
    y2, i2 = oracle.decompose(x * 2.0**7)
    assert i1["windows"].tolist() == i2["windows"].tolist()
    assert np.array_equal(y1 * 2.0**7, y2)


def test_constant_input_has_no_mode():
    y, info = oracle.decompose(np.full(256, 1.5))
    assert info["n_imfs"].tolist() == [0] and y.shape == (1, 1, 256)
    assert np.array_equal(y[0, 0], np.full(256, 1.5))


def test_non_finite_input_fails_with_code_1():
    x = np.sin(np.arange(300.0) / 3)
    x[10] = np.nan
    y, info = oracle.decompose(np.stack([x, np.sin(np.arange(300.0) / 5)]), dim=1)
    assert info["error"].tolist()[0] == 1 and np.isnan(y[0]).all()
    assert info["error"].tolist()[1] == 0


def test_max_imf_is_a_prefix():
    x = np.load(GOLDEN / "corpus" / "noiseflat_512.npz")["x"]
    y, info = oracle.decompose(x)
    y2, info2 = oracle.decompose(x, max_imf=2)
    assert info2["n_imfs"].tolist() == [2]
    assert np.array_equal(y2[0, :2], y[0, :2])


# ---- the cube filter -------------------------------------------------------------------------------
def _reflect_pad(x, r):
    return np.concatenate([x[:r][::-1], x, x[len(x) - r:][::-1]]) if r else x


def brute(x, w, op):
    n = x.size
    r = min((w - 1) // 2, n - 1)
    e = _reflect_pad(x, r)
    win = np.lib.stride_tricks.sliding_window_view(e, 2 * r + 1)
    return {"max": win.max(axis=1), "min": win.min(axis=1), "sum": win.sum(axis=1)}[op]


@pytest.mark.parametrize("n", [1, 2, 5, 31, 32, 33, 100, 517])
@pytest.mark.parametrize("w", [1, 3, 7, 33, 65, 101, 1025])
def test_vhgw_line_matches_the_sliding_window(n, w):
    rng = np.random.default_rng(n * 1000 + w)
    x = rng.standard_normal(n)
    for op in ("max", "min"):
        assert np.array_equal(oracle.vhgw_line(x, w, op), brute(x, w, op))
    got = oracle.vhgw_line(x, w, "sum")
    ref = brute(x, w, "sum")
    assert np.allclose(got, ref, rtol=0, atol=1e-12 * max(1.0, np.abs(x).sum()))


def test_vhgw_sum_aligned_window_takes_the_mean_of_both_folds():
    # n = w: the one segment starting at 0 is the whole extended window of the middle sample
    x = np.arange(1.0, 8.0)
    got = oracle.vhgw_line(x, 7, "sum")
    assert np.allclose(got, brute(x, 7, "sum"))


@pytest.mark.parametrize("shape", [(40,), (17, 23), (6, 7, 8)])
def test_cube_is_the_separable_reflect_filter(shape):
    rng = np.random.default_rng(3)
    x = rng.standard_normal(shape)
    for w in (3, 9):
        assert np.array_equal(oracle.cube(x, "max", w), scipy.ndimage.maximum_filter(x, size=w, mode="reflect")
                              if w <= 2 * min(shape) - 1 else oracle.cube(x, "max", w))
        assert np.array_equal(oracle.cube(x, "min", w), scipy.ndimage.minimum_filter(x, size=w, mode="reflect"))
        assert np.allclose(oracle.cube(x, "box", w), scipy.ndimage.uniform_filter(x, size=w, mode="reflect"),
                           rtol=1e-12, atol=1e-12)


def test_cube_clamps_the_window_per_axis():
    x = np.random.default_rng(4).standard_normal((5, 40))
    got = oracle.cube(x, "box", 21)  # axis 0 clamps to 2 * 4 + 1 = 9
    ref = scipy.ndimage.uniform_filter(x, size=(9, 21), mode="reflect")
    assert np.allclose(got, ref, rtol=1e-12, atol=1e-12)


# ---- selections and noise --------------------------------------------------------------------------
def test_quantile_matches_the_numpy_table():
    table = json.loads((GOLDEN / "quantile_table.json").read_text())
    for row in table["rows"]:
        v = [i * i + 0.5 for i in range(row["n"])]
        for q, want in zip(table["q"], row["values"]):
            assert oracle.quantile(v, q) == want
    assert oracle.quantile([1.0, 2.0, 10.0, 11.0], 0.25) == 1.75


@pytest.mark.parametrize("n", [1, 2, 3, 10, 1001])
def test_median_and_quantile_are_numpy(n):
    v = np.random.default_rng(n).standard_normal(n)
    assert oracle.median(v) == np.median(v)
    for q in (0.0, 0.1, 0.25, 0.5, 0.9, 1.0):
        assert oracle.quantile(v, q) == np.quantile(v, q)
    assert np.isnan(oracle.median([]))


def test_noise_sigma_closed_form():
    rng = np.random.default_rng(5)
    r = rng.standard_normal((30, 40))
    d = np.concatenate([np.diff(r, axis=0).ravel(), np.diff(r, axis=1).ravel()])
    mad = np.median(np.abs(d - np.median(d)))
    assert oracle.noise_sigma(r, False) == (1.4826 * mad) / np.sqrt(2.0)
    per = []
    for a in range(2):
        da = np.diff(r, axis=a).ravel()
        per.append((1.4826 * np.median(np.abs(da - np.median(da)))) / np.sqrt(2.0))
    assert oracle.noise_sigma(r, True) == max(per)


# ---- extrema ------------------------------------------------------------------------------------------
def _scipy_flags(x, tau):
    out = np.zeros(x.size, bool)
    for sgn in (1, -1):
        peaks, _ = scipy.signal.find_peaks(sgn * x)
        if peaks.size:
            prom = scipy.signal.peak_prominences(sgn * x, peaks)[0]
            out[peaks[prom >= tau[peaks]]] = True
    return out


@pytest.mark.parametrize("seed", range(6))
def test_sig_flags_are_find_peaks_with_prominence(seed):
    rng = np.random.default_rng(seed)
    x = np.round(rng.standard_normal(400) * (3 if seed % 2 else 50)) / 3  # plateaus when coarse
    tau = np.full(400, 0.4)
    flags, ext = oracle.sig_flags(x, tau)
    ref = _scipy_flags(x, tau)
    assert np.array_equal(flags.astype(bool), ref)
    assert int(ext[0]) == int(ref.sum())


def test_sig_flags_2d_axes():
    rng = np.random.default_rng(9)
    x = rng.standard_normal((20, 30))
    tau = np.full_like(x, 0.3)
    flags, ext = oracle.sig_flags(x, tau)
    for i in range(20):
        assert np.array_equal((flags[i] >> 1) & 1, _scipy_flags(x[i], tau[i]))
    for j in range(30):
        assert np.array_equal(flags[:, j] & 1, _scipy_flags(x[:, j], tau[:, j]))


# ---- components ----------------------------------------------------------------------------------------
@pytest.mark.parametrize("shape", [(200,), (30, 40), (10, 11, 12)])
def test_ccl_is_scipy_label_canonically_relabelled(shape):
    m = np.random.default_rng(1).random(shape) < 0.5
    lab = oracle.ccl(m.astype(np.uint8))
    ref, n = scipy.ndimage.label(m)
    flat_ref = ref.ravel()
    canon = np.full(n + 1, -1)
    for comp in range(1, n + 1):
        canon[comp] = np.flatnonzero(flat_ref == comp).min()
    assert np.array_equal(lab.ravel(), np.where(flat_ref > 0, canon[flat_ref], -1))


# ---- DCT, Riesz, eigen ------------------------------------------------------------------------------------
@pytest.mark.parametrize("shape", [(50,), (12, 20), (6, 8, 10)])
def test_dct_is_the_orthonormal_dctn(shape):
    x = np.random.default_rng(2).standard_normal(shape)
    X = oracle.dct(x)
    assert np.allclose(X, scipy.fft.dctn(x, norm="ortho"), atol=1e-12)
    assert np.allclose(oracle.dct(X, inverse=True), x, atol=1e-12)


def test_inverse_sin_basis_is_the_dst_of_the_shifted_spectrum():
    # sum_k c_k Y(k) sin(pi k (2x+1) / (2n)) = DST-III of Y(1..n-1) with a zero at n
    n = 16
    Y = np.random.default_rng(3).standard_normal(n)
    got = oracle.dct(Y, inverse=True, odd=1)
    z = np.concatenate([Y[1:], [0.0]])
    ref = scipy.fft.idst(z, type=2, norm="ortho")
    assert np.allclose(got, ref, atol=1e-12)


@pytest.mark.parametrize("k0,k1,w", [(10, 6, 9), (3, 17, 5), (0, 8, 16), (12, 0, 11)])
def test_riesz_of_a_dct_basis_field(k0, k1, w):
    # A single DCT-II basis field f = cos(th0) cos(th1), th_a = pi k_a (2 x_a + 1) / (2 n): its band
    # is G f with G the log-Gabor at rad = |u|, u_a = k_a / (2 n), and the Riesz components are
    # R1_0 = G (u0/rad) sin(th0) cos(th1), R1_1 = G (u1/rad) cos(th0) sin(th1),
    # R2_00 = -G (u0/rad)^2 f, R2_01 = G (u0 u1 / rad^2) sin(th0) sin(th1), R2_11 = -G (u1/rad)^2 f.
    n = 48
    y, x = np.meshgrid(np.arange(n), np.arange(n), indexing="ij")
    th0 = np.pi * k0 * (2 * y + 1) / (2 * n)
    th1 = np.pi * k1 * (2 * x + 1) / (2 * n)
    f = np.cos(th0) * np.cos(th1)
    u = np.array([k0, k1]) / (2 * n)
    rad = np.linalg.norm(u)
    l2 = np.log(2.0)
    sg = l2 / (2 * np.sqrt(2 * l2))
    G = np.exp(-np.log(rad * w) ** 2 / (2 * sg * sg))
    r1, r2 = oracle.riesz(f, w)
    a, b = u / rad
    tol = 1e-12
    assert np.allclose(r1[0], G * a * np.sin(th0) * np.cos(th1), atol=tol)
    assert np.allclose(r1[1], G * b * np.cos(th0) * np.sin(th1), atol=tol)
    assert np.allclose(r2[0], -G * a * a * f, atol=tol)
    assert np.allclose(r2[1], G * a * b * np.sin(th0) * np.sin(th1), atol=tol)
    assert np.allclose(r2[2], -G * b * b * f, atol=tol)


def test_eigen_matches_eigh():
    rng = np.random.default_rng(4)
    for D in (2, 3):
        K = D * (D + 1) // 2
        A = rng.standard_normal((200, D, D))
        S = A @ A.transpose(0, 2, 1)
        iu = np.triu_indices(D)
        t = S[:, iu[0], iu[1]]
        nv, c = oracle.eigen(t, D)
        w, V = np.linalg.eigh(S)
        top = V[:, :, -1]
        assert np.allclose(np.abs((nv * top).sum(axis=1)), 1.0, atol=1e-10)
        ref_c = np.clip((w[:, -1] - w[:, -2]) / (w[:, -1] + w[:, -2]), 0, 1)
        assert np.allclose(c, ref_c, atol=1e-12)
        assert K == t.shape[1]


# ---- the element --------------------------------------------------------------------------------------------
def test_rung_ladder_matches_the_kat_table():
    rows = json.loads((GOLDEN / "rungs_table.json").read_text())
    assert len(rows) == 66
    for row in rows:
        got = oracle.rung_ladder(float.fromhex(row["d_min"]), float.fromhex(row["d_max"]), row["w"], row["n_max"])
        assert got == row["rungs"], row


def test_interp_is_map_coordinates_on_the_reflected_extension():
    rng = np.random.default_rng(6)
    f = rng.standard_normal((9, 11))
    pts = rng.uniform(-12, 24, (500, 2))
    got = oracle.interp(f, pts)
    # the periodic half-sample reflection of period 2n per axis = map_coordinates(mode='grid-mirror')
    ref = scipy.ndimage.map_coordinates(f, pts.T, order=1, mode="grid-mirror")
    assert np.allclose(got, ref, atol=1e-12)
