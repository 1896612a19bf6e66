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


# ----------------------------------------------------------------------------------------------
# The monogenic signal (reference src/eda/_api.py:356-426, cpp/src/method/orientation.cu:420-493)

MONO_EPS = {np.float32: float(np.finfo(np.float32).eps), np.float64: float(np.finfo(np.float64).eps)}
MONO_ZERO = {np.float32: 2.0**-14, np.float64: 2.0**-30}


def _wrap(d):
    return np.abs(np.angle(np.exp(1j * np.asarray(d, dtype=np.float64))))


def _dct_basis(n, m):
    """cos(pi m (x + 1/2) / n) and its angle reduced exactly to [0, 2 pi)."""
    theta = np.pi * ((m * (2 * np.arange(n) + 1)) % (4 * n)) / (2 * n)
    return np.cos(theta), theta


def _waves(shape, seed, offset=0.3):
    rng = np.random.default_rng(seed)
    g = np.meshgrid(*[np.arange(n, dtype=np.float64) for n in shape], indexing="ij")
    out = np.full(shape, float(offset))
    for period in (5.0, 9.0, 23.0):
        k = rng.normal(size=len(shape))
        k = k / np.linalg.norm(k) * 2 * np.pi / period
        out += rng.uniform(0.5, 1.0) * np.sin(sum(ki * gi for ki, gi in zip(k, g)) + rng.uniform(0, 6))
    return out


def _fft_monogenic(f, zero):
    """The same definitions through the FFT of the half-sample symmetric extension (2 n per axis),
    float64, an algorithm independent of the oracle's DCT sums: R_a has the multiplier -i u_a / |u|,
    d/dx_a / (2 pi) i u_a, D_b R_a u_a u_b / |u|, R_a R_b -u_a u_b / |u|^2 (u in cycles per
    sample, every multiplier 0 at the DC). Returns the outputs and (mu1, mu2) of T."""
    f = np.asarray(f, np.float64)
    D = f.ndim
    ext = f
    for a in range(D):
        ext = np.concatenate([ext, np.flip(ext, axis=a)], axis=a)
    F = np.fft.fftn(ext)
    u = np.meshgrid(*[np.fft.fftfreq(2 * n) for n in f.shape], indexing="ij")
    rad = np.sqrt(sum(v * v for v in u))
    inv = np.divide(1.0, rad, out=np.zeros_like(rad), where=rad > 0)
    crop = tuple(slice(0, n) for n in f.shape)

    def back(mult):
        return np.real(np.fft.ifftn(F * mult))[crop]

    r1 = np.stack([back(-1j * u[a] * inv) for a in range(D)])
    df = np.stack([back(1j * u[a]) for a in range(D)])
    pairs = [(a, b) for a in range(D) for b in range(D)]
    dr = {(a, b): back(u[a] * u[b] * inv) for a, b in pairs}
    out = {"mu": None}
    if D == 1:
        n = np.ones((1,) + f.shape)
    else:
        h = np.zeros(f.shape + (D, D))
        for a, b in pairs:
            h[..., a, b] = back(-u[a] * u[b] * inv**2)
        T = np.einsum("a...,b...->...ab", r1, r1) + h @ h
        w, V = np.linalg.eigh(T)
        n = np.moveaxis(V[..., :, -1], -1, 0)
        sign = np.zeros(f.shape)
        for a in range(D):
            sign = np.where((sign == 0) & (np.abs(n[a]) > zero), np.sign(n[a]), sign)
        n = n * np.where(sign < 0, -1.0, 1.0)
        s = w[..., -1] + w[..., -2]
        c = np.divide(w[..., -1] - w[..., -2], s, out=np.zeros_like(s), where=s > 0)
        out.update(direction=n, coherence=np.clip(c, 0, 1), mu=(w[..., -1], w[..., -2]))
    q = (n * r1).sum(axis=0)
    a2 = f * f + (r1 * r1).sum(axis=0)
    dnq = sum(n[a] * n[b] * dr[(a, b)] for a, b in pairs)
    dnf = (n * df).sum(axis=0)
    out.update(amplitude=np.sqrt(a2), phase=np.arctan2(q, f), frequency=(f * dnq - q * dnf) / a2)
    return out


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
@pytest.mark.parametrize(("n", "m"), [(16, 1), (37, 5), (64, 21), (255, 254), (1000, 3)])
def test_monogenic_of_a_dct_basis_field_1d(n, m, dtype):
    """cos(pi m (x + 1/2) / n) is a plane wave of frequency m / (2 n) on the extension: amplitude 1,
    phase pi m (x + 1/2) / n and frequency m / (2 n) at every sample."""
    f, theta = _dct_basis(n, m)
    r = oracle.monogenic(f.astype(dtype))
    eps = MONO_EPS[dtype]
    assert r["direction"] is None and r["coherence"] is None
    assert np.abs(r["amplitude"] - 1).max() <= 64 * eps
    assert np.abs(r["frequency"] - m / (2 * n)).max() <= 24 * eps
    assert _wrap(r["phase"] - theta).max() <= 48 * eps


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
@pytest.mark.parametrize("shape", [(12, 20), (6, 9, 10)])
def test_monogenic_of_a_dct_basis_field_along_one_axis(shape, dtype):
    """The basis field along axis a of a 2-D or 3-D field: the same values, direction the unit vector
    of axis a, coherence 1."""
    eps = MONO_EPS[dtype]
    for axis, n in enumerate(shape):
        m = n // 3 + 1
        f, theta = _dct_basis(n, m)
        view = [1] * len(shape)
        view[axis] = n
        r = oracle.monogenic(np.broadcast_to(f.reshape(view), shape).astype(dtype))
        assert np.abs(r["amplitude"] - 1).max() <= 64 * eps
        assert np.abs(r["frequency"] - m / (2 * n)).max() <= 24 * eps
        assert _wrap(r["phase"] - np.broadcast_to(theta.reshape(view), shape)).max() <= 48 * eps
        unit = np.zeros((len(shape),) + (1,) * len(shape))
        unit[axis] = 1
        assert np.abs(r["direction"] - unit).max() <= 8 * eps
        assert np.abs(r["coherence"] - 1).max() <= 4 * eps


@pytest.mark.parametrize("seed", [0, 1])
@pytest.mark.parametrize("dtype", [np.float64, np.float32])
@pytest.mark.parametrize("shape", [(300,), (37, 45), (17, 20, 23)], ids=["1d", "2d", "3d"])
def test_monogenic_matches_the_fft_computation(shape, dtype, seed):
    """Against the FFT of the extension in float64, in eps of the field's scale: the amplitude, the
    phase weighted by A, the frequency by A^2, n by the eigenvalue gap and c by mu1 + mu2 (each
    output is only as well defined as its weight)."""
    x = _waves(shape, seed).astype(dtype)
    r = oracle.monogenic(x)
    ref = _fft_monogenic(x, MONO_ZERO[dtype])
    eps, scale = MONO_EPS[dtype], float(np.abs(x).max())
    A = ref["amplitude"]
    errors = {
        "amplitude": np.abs(r["amplitude"] - A).max() / scale,
        "phase": (_wrap(r["phase"] - ref["phase"]) * A).max() / scale,
        "frequency": (np.abs(r["frequency"] - ref["frequency"]) * A**2).max() / scale**2,
    }
    if ref["mu"] is not None:
        mu1, mu2 = ref["mu"]
        dn = np.abs(r["direction"] - ref["direction"]).max(axis=0)
        errors["direction"] = (dn * (mu1 - mu2)).max() / mu1.max()
        errors["coherence"] = (np.abs(r["coherence"] - ref["coherence"]) * (mu1 + mu2)).max() / mu1.max()
    tol = {"amplitude": 32, "phase": 512, "frequency": 16, "direction": 32, "coherence": 64}
    for name, value in errors.items():
        assert value / eps <= tol[name], (name, value / eps)


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
@pytest.mark.parametrize("shape", [(50,), (9, 13), (5, 6, 7)])
def test_monogenic_of_constant_and_zero_fields(shape, dtype):
    for value in (0.0, 0.75, -32.0):
        r = oracle.monogenic(np.full(shape, value, dtype=dtype))
        assert (r["amplitude"] == abs(value)).all() and (r["frequency"] == 0).all()
        assert (r["phase"] == (np.asarray(np.pi, dtype) if value < 0 else 0)).all()
        if len(shape) > 1:
            assert (r["direction"][0] == 1).all() and (r["direction"][1:] == 0).all()
            assert (r["coherence"] == 0).all()


@pytest.mark.parametrize("j", [-8, 3, 20])
def test_monogenic_gain_by_a_power_of_two(j):
    """The field is prescaled to max |f| in [1, 2): a gain 2^j scales the amplitude and leaves every
    other output's bits."""
    x = _waves((30, 28), 8)
    r = oracle.monogenic(x)
    g = oracle.monogenic(x * 2.0**j)
    assert np.array_equal(g["amplitude"], r["amplitude"] * 2.0**j)
    for name in ("phase", "frequency", "direction", "coherence"):
        assert np.array_equal(g[name], r[name])
