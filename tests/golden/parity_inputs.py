# SPDX-License-Identifier: Apache-2.0
"""The inputs of the reference's parity manifest (tests/data/parity_v2.json: 287 cases over 147
recipes), regenerated bit for bit (test infrastructure).

This restates the input algorithm of the reference's tests/python/cases.py:1-741. The bytes of an
input are fixed by its steps, so the steps here are the reference's, in the same order: random
numbers are the raw 64-bit words of NumPy's PCG64 (seeded by a SeedSequence of the recipe name's
SHA-256 and the element number), mapped to [0, 1) as (word >> 11) 2^-53; the fields use only IEEE
elementwise arithmetic and a polynomial cosine, sine and exp(-t) (Taylor series after an exact
range reduction), so no libm or NumPy transcendental decides a bit. Each manifest entry's
input_sha256 checks the restatement.

Layout: a recipe builds each batch element's field (float64, spatial axes in the order of `axes`),
transposes it into the array, optionally rounds it down to a quantum, scales it (2^k, a gain) and
shifts it, casts it to the case's dtype and then writes its special elements (a constant 0.75
element, one NaN or one +inf sample two fifths into the element).
"""

from __future__ import annotations

import hashlib
import math

import numpy as np

# ---- polynomial elementary functions (float64) ---------------------------------------------------
TWO_PI = 6.283185307179586            # the double nearest 2 pi
TWO_PI_TAIL = 2.4492935982947064e-16  # 2 pi - TWO_PI
INV_TWO_PI = 0.15915494309189535
LN2_HEAD = 0.6931471805599453
LN2_TAIL = 2.3190468138462996e-17
INV_LN2 = 1.4426950408889634
# Taylor coefficients as correctly rounded quotients of integers
COS_TERMS = tuple((-1) ** k / math.factorial(2 * k) for k in range(16))
SIN_TERMS = tuple((-1) ** k / math.factorial(2 * k + 1) for k in range(16))
EXP_TERMS = tuple((-1) ** k / math.factorial(k) for k in range(23))


def trunc_floor(t: np.ndarray) -> np.ndarray:
    """floor(t), |t| < 2^62: truncation towards zero, minus one where that rounded up."""
    whole = t.astype(np.int64).astype(np.float64)
    return whole - (whole > t).astype(np.float64)


def poly(terms: tuple[float, ...], z: np.ndarray) -> np.ndarray:
    """Horner's rule from the highest term down."""
    acc = np.full(z.shape, terms[-1])
    for c in reversed(terms[:-1]):
        acc = acc * z + c
    return acc


def to_pi(x: np.ndarray) -> np.ndarray:
    """x minus the nearest multiple of 2 pi (two-part constant)."""
    k = trunc_floor(x * INV_TWO_PI + 0.5)
    return (x - k * TWO_PI) - k * TWO_PI_TAIL


def cos_p(x) -> np.ndarray:
    r = to_pi(np.asarray(x, dtype=np.float64))
    return poly(COS_TERMS, r * r)


def sin_p(x) -> np.ndarray:
    r = to_pi(np.asarray(x, dtype=np.float64))
    return r * poly(SIN_TERMS, r * r)


def exp_neg(t) -> np.ndarray:
    """exp(-t), t >= 0: t = k ln 2 + r, exp(-r) 2^-k."""
    t = np.minimum(np.asarray(t, dtype=np.float64), 1400.0)
    k = trunc_floor(t * INV_LN2)
    r = (t - k * LN2_HEAD) - k * LN2_TAIL
    return np.ldexp(poly(EXP_TERMS, r), -k.astype(np.int32))


# ---- random numbers --------------------------------------------------------------------------------
class Draws:
    """Doubles in [0, 1) from the raw words of a PCG64 generator keyed by (seed, element)."""

    def __init__(self, *key: int) -> None:
        self.gen = np.random.PCG64(np.random.SeedSequence(list(key)))

    def uniform(self, shape: tuple[int, ...]) -> np.ndarray:
        size = math.prod(shape)
        raw = self.gen.random_raw(size) if size else np.zeros(0, np.uint64)
        top = (np.asarray(raw, dtype=np.uint64) >> np.uint64(11)).astype(np.float64)
        return (top * 2.0**-53).reshape(shape)

    def one(self) -> float:
        return float(self.uniform((1,))[0])

    def within(self, lo: float, hi: float) -> float:
        return lo + (hi - lo) * self.one()

    def gauss(self, shape: tuple[int, ...]) -> np.ndarray:
        """Unit variance: the sum of four uniforms, minus 2, times sqrt 3."""
        acc = self.uniform(shape)
        for _ in range(3):
            acc = acc + self.uniform(shape)
        return (acc - 2.0) * math.sqrt(3.0)

    def unit(self, dim: int) -> list[float]:
        """A unit vector from normalized gauss draws (drawn again while nearly zero)."""
        while True:
            v = [float(e) for e in self.gauss((dim,))]
            sq = 0.0
            for e in v:
                sq += e * e
            if sq > 1e-6:
                length = math.sqrt(sq)
                return [e / length for e in v]


def name_seed(name: str) -> int:
    return int.from_bytes(hashlib.sha256(name.encode()).digest()[:8], "little")


# ---- fields: family(n, draws, **params) -> float64 array of shape n ----------------------------------
def grid(n: tuple[int, ...]) -> list[np.ndarray]:
    """The coordinate of every axis, shaped to broadcast over n."""
    axes = []
    for a, m in enumerate(n):
        shape = [1] * len(n)
        shape[a] = m
        axes.append(np.arange(m, dtype=np.int64).astype(np.float64).reshape(shape))
    return axes


def along(d: list[float], c: list[np.ndarray], n: tuple[int, ...]) -> np.ndarray:
    acc = np.zeros(n)
    for a in range(len(n)):
        acc = acc + c[a] * d[a]
    return acc


def sq_dist(center: list[float], c: list[np.ndarray], n: tuple[int, ...]) -> np.ndarray:
    acc = np.zeros(n)
    for a in range(len(n)):
        e = c[a] - center[a]
        acc = acc + e * e
    return acc


def plane_wave(n, c, u: Draws, period: float, amp: float = 1.0) -> np.ndarray:
    d = u.unit(len(n))
    phi = TWO_PI * u.one()
    return amp * cos_p(along(d, c, n) * (TWO_PI / period) + phi)


def fam_constant(n, u, value=0.75):
    return np.full(n, float(value))


def fam_noise(n, u, smooth=0):
    v = u.gauss(n)
    for _ in range(smooth):
        for a in range(len(n)):
            v = (np.roll(v, -1, axis=a) + v + np.roll(v, 1, axis=a)) / 3.0
    return v


def fam_tones(n, u, count=3, base=5.0, ratio=3.0):
    c = grid(n)
    v = np.zeros(n)
    for j in range(count):
        period = base * ratio**j * u.within(0.8, 1.2)
        amp = u.within(0.5, 1.5)  # drawn before the wave's own draws
        v = v + plane_wave(n, c, u, period, amp)
    return v


def fam_chirp(n, u, f0=0.01, f1=0.3):
    c = grid(n)
    d = u.unit(len(n))
    t = along(d, c, n)
    span = float(max(n))
    lo = f0 * u.within(0.5, 1.5)
    hi = f1 * u.within(0.85, 1.15)
    phase = TWO_PI * (lo * t + 0.5 * (hi - lo) * t * t / span) + TWO_PI * u.one()
    return cos_p(phase) + 0.3 * plane_wave(n, c, u, 0.6 * span)


def fam_amfm(n, u, period=7.0):
    c = grid(n)
    d = u.unit(len(n))
    t = along(d, c, n)
    p = period * u.within(0.8, 1.3)
    am = 1.0 + 0.6 * cos_p(t * (TWO_PI / (8.0 * p)))
    fm = 2.0 * sin_p(t * (TWO_PI / (5.0 * p)))
    carrier = am * cos_p(t * (TWO_PI / p) + fm + TWO_PI * u.one())
    return carrier + plane_wave(n, c, u, 25.0 * p, 0.5)


def fam_blobs(n, u, count=6, size=(0.05, 0.17)):
    c = grid(n)
    v = np.zeros(n)
    small = float(min(n))
    for _ in range(count):
        center = [u.within(0.0, float(m)) for m in n]
        s = max(1.0, small * u.within(size[0], size[1]))
        amp = u.within(0.5, 1.5)
        amp = amp * (1.0 if u.one() < 0.6 else -1.0)
        v = v + amp * exp_neg(sq_dist(center, c, n) / (2.0 * s * s))
    return v


def fam_fringes(n, u):
    c = grid(n)
    first = plane_wave(n, c, u, u.within(4.0, 7.0))
    return first + plane_wave(n, c, u, u.within(14.0, 24.0), 0.7)


def fam_gabor(n, u, count=3, background=0.3, p_lo=3.0, p_hi=8.0, width=(1.5, 3.0)):
    c = grid(n)
    v = np.zeros(n)
    for _ in range(count):
        center = [u.within(0.15, 0.85) * m for m in n]
        p = u.within(p_lo, p_hi)
        s = u.within(width[0], width[1]) * p
        d = u.unit(len(n))
        t = along(d, c, n)
        for a in range(len(n)):
            t = t + (-center[a]) * d[a]
        env = exp_neg(sq_dist(center, c, n) / (2.0 * s * s))
        weight = u.within(0.6, 1.4)
        v = v + weight * env * cos_p(t * (TWO_PI / p) + TWO_PI * u.one())
    return v + plane_wave(n, c, u, 0.8 * float(max(n)), background)


def fam_edges(n, u, count=4, noise=0.02):
    c = grid(n)
    v = np.zeros(n)
    for _ in range(count):
        d = u.unit(len(n))
        cut = 0.0
        for a in range(len(n)):
            cut += d[a] * u.within(0.2, 0.8) * n[a]
        step = (along(d, c, n) > cut).astype(np.float64)
        v = v + u.within(-1.0, 1.0) * step
    return v + noise * u.gauss(n)


def fam_flat(n, u, period=5.0):
    c = grid(n)
    inside = np.ones(n, dtype=bool)
    for a, m in enumerate(n):
        lo = u.within(0.1, 0.35) * m
        hi = u.within(0.6, 0.9) * m
        inside = inside & (c[a] >= lo) & (c[a] < hi)
    tone = plane_wave(n, c, u, period * u.within(0.8, 1.2))
    d = u.unit(len(n))
    mid = 0.0
    for a, m in enumerate(n):
        mid += 0.5 * d[a] * m
    step = (along(d, c, n) > mid).astype(np.float64)
    return tone * inside.astype(np.float64) + 0.5 * step


def fam_local(n, u, fine=5.0, frac=(0.35, 0.65), coarse_period=0.45):
    c = grid(n)
    coarse = plane_wave(n, c, u, coarse_period * float(max(n)) * u.within(0.8, 1.2))
    cut = u.within(frac[0], frac[1]) * n[0]
    part = (c[0] < cut).astype(np.float64) * np.ones(n)
    return coarse + part * plane_wave(n, c, u, fine * u.within(0.8, 1.2), 0.7)


def fam_growth(n, u, frac=(0.55, 0.62), fine=4.5, noise=0.02):
    c = grid(n)
    p = u.within(frac[0], frac[1]) * n[0]
    coarse = cos_p(c[0] * (TWO_PI / p) + TWO_PI * u.one()) * np.ones(n)
    fine_wave = plane_wave(n, c, u, fine * u.within(0.9, 1.1), 0.4)
    return fine_wave + coarse + noise * u.gauss(n)


def fam_zone(n, u, top=0.35):
    c = grid(n)
    center = [0.5 * m + u.within(-1.0, 1.0) for m in n]
    rim = 0.0
    for m in n:
        rim += 0.25 * m * m
    rim = math.sqrt(rim)
    return cos_p(sq_dist(center, c, n) * (TWO_PI * 0.5 * top / rim) + TWO_PI * u.one())


def fam_mixture(n, u):
    c = grid(n)
    d = u.unit(len(n))
    trend = along(d, c, n) * (0.5 / float(max(n)))
    return fam_tones(n, u, count=2) + fam_blobs(n, u, count=2) + 0.15 * u.gauss(n) + trend


FAMILIES = {
    "constant": fam_constant, "noise": fam_noise, "tones": fam_tones, "chirp": fam_chirp,
    "amfm": fam_amfm, "blobs": fam_blobs, "fringes": fam_fringes, "gabor": fam_gabor,
    "edges": fam_edges, "flat": fam_flat, "local": fam_local, "zone": fam_zone,
    "growth": fam_growth, "mixture": fam_mixture,
}

# ---- the recipe table (the reference's tests/python/cases.py:367-688) --------------------------------
# (name, family, shape, options). Options: axes, params, dtypes, gain, gain_pow2, offset, nan_at,
# inf_at, constant_at, seed_from, quantum (the call arguments are in the manifest's entries).
_F32, _F64 = ("float32",), ("float64",)
_MANY = {"count": 40, "size": (0.04, 0.1)}
_FINE = {"frac": (0.15, 0.3), "fine": 4.0}
RECIPES = [
    ("d1-constant-2", "constant", (2,), {}),
    ("d1-noise-2", "noise", (2,), {}),
    ("d1-noise-3", "noise", (3,), {}),
    ("d1-tones-3", "tones", (3,), {}),
    ("d1-tones-37", "tones", (37,), {}),
    ("d1-noise-37", "noise", (37,), {}),
    ("d1-chirp-37", "chirp", (37,), {}),
    ("d1-tones-512", "tones", (512,), {}),
    ("d1-mixture-512", "mixture", (512,), {}),
    ("d1-chirp-512", "chirp", (512,), {}),
    ("d1-amfm-512", "amfm", (512,), {}),
    ("d1-noise-512", "noise", (512,), {}),
    ("d1-rednoise-512", "noise", (512,), {"params": {"smooth": 3}}),
    ("d1-edges-512", "edges", (512,), {}),
    ("d1-flat-512", "flat", (512,), {}),
    ("d1-gabor-512", "gabor", (512,), {}),
    ("d1-local-512", "local", (512,), {}),
    ("d1-blobs-512", "blobs", (512,), {}),
    ("d1-constant-512", "constant", (512,), {}),
    ("d1-tones-2048", "tones", (2048,), {}),
    ("d1-mixture-2048", "mixture", (2048,), {}),
    ("d1-chirp-2048", "chirp", (2048,), {}),
    ("d1-amfm-2048", "amfm", (2048,), {}),
    ("d1-gabor-2048", "gabor", (2048,), {}),
    ("d1-local-2048", "local", (2048,), {}),
    ("d1-noise-4096", "noise", (4096,), {}),
    ("d1-mixture-4096", "mixture", (4096,), {}),
    ("d1-local-4096", "local", (4096,), {}),
    ("d1-chirp-8192", "chirp", (8192,), {"dtypes": _F32}),
    ("d1-mixture-8192", "mixture", (8192,), {"dtypes": _F32}),
    ("d1-batch4-512", "mixture", (4, 512), {"axes": (1,)}),
    ("d1-chunk1-512", "gabor", (5, 512), {"axes": (1,)}),
    ("d1-chunk2-512", "gabor", (5, 512), {"axes": (1,)}),
    ("d1-chunk5-512", "gabor", (5, 512), {"axes": (1,)}),
    ("d1-twobatch-37", "tones", (2, 37, 3), {"axes": (1,)}),
    ("d1-batchfirst-axis0", "chirp", (512, 3), {"axes": (0,)}),
    ("d1-nan-status", "mixture", (3, 512), {"axes": (1,), "nan_at": (1,)}),
    ("d1-constant-element", "mixture", (3, 512), {"axes": (1,), "constant_at": (2,)}),
    ("d1-inf-nocheck", "mixture", (2, 512), {"axes": (1,), "inf_at": (0,)}),
    ("d1-nan-nocheck", "mixture", (2, 512), {"axes": (1,), "nan_at": (1,)}),
    ("d1-inf-check", "mixture", (2, 512), {"axes": (1,), "inf_at": (1,)}),
    ("d1-gain-p40", "mixture", (512,), {"gain_pow2": 40, "seed_from": "d1-mixture-512"}),
    ("d1-gain-m40", "mixture", (512,), {"gain_pow2": -40, "seed_from": "d1-mixture-512"}),
    ("d1-gain-1e30", "mixture", (512,), {"gain": 1e30, "seed_from": "d1-mixture-512"}),
    ("d1-gain-1e-30", "mixture", (512,), {"gain": 1e-30, "seed_from": "d1-mixture-512"}),
    ("d1-gain-1e300", "mixture", (512,), {"gain": 1e300, "dtypes": _F64, "seed_from": "d1-mixture-512"}),
    ("d1-gain-1e-300", "mixture", (512,), {"gain": 1e-300, "dtypes": _F64, "seed_from": "d1-mixture-512"}),
    ("d1-negative", "mixture", (512,), {"gain": -1.0, "seed_from": "d1-mixture-512"}),
    ("d1-negative-3", "gabor", (2048,), {"gain": -3.0}),
    ("d1-offset", "mixture", (512,), {"offset": 100.0, "seed_from": "d1-mixture-512"}),
    ("d1-maximf1", "mixture", (2048,), {}),
    ("d1-maximf2", "mixture", (2048,), {}),
    ("d1-tie0", "mixture", (512,), {"quantum": 0.0625}),
    ("d1-quantized", "mixture", (512,), {"quantum": 0.0625, "seed_from": "d1-tie0"}),
    ("d1-tie2m20", "mixture", (512,), {"quantum": 2.0**-12}),
    ("d1-tie1e-3", "edges", (512,), {}),
    ("d1-tie1e-3-quant", "edges", (512,), {"params": {"noise": 0.0}}),
    ("d2-noise-2x16", "noise", (2, 16), {}),
    ("d2-fringes-2x16", "fringes", (2, 16), {}),
    ("d2-constant-16x12", "constant", (16, 12), {}),
    ("d2-noise-16x12", "noise", (16, 12), {}),
    ("d2-fringes-16x12", "fringes", (16, 12), {}),
    ("d2-tones-24x31", "tones", (24, 31), {}),
    ("d2-blobs-24x31", "blobs", (24, 31), {"params": _MANY}),
    ("d2-edges-24x31", "edges", (24, 31), {}),
    ("d2-fringes-40x36", "fringes", (40, 36), {}),
    ("d2-gabor-40x36", "gabor", (40, 36), {}),
    ("d2-mixture-40x36", "mixture", (40, 36), {}),
    ("d2-flat-40x36", "flat", (40, 36), {}),
    ("d2-zone-64x64", "zone", (64, 64), {}),
    ("d2-gabor-64x64", "gabor", (64, 64), {}),
    ("d2-amfm-64x64", "amfm", (64, 64), {}),
    ("d2-local-64x64", "local", (64, 64), {}),
    ("d2-edges-64x64", "edges", (64, 64), {}),
    ("d2-blobs-64x64", "blobs", (64, 64), {"params": _MANY}),
    ("d2-flat-64x64", "flat", (64, 64), {}),
    ("d2-fringes-128x128", "fringes", (128, 128), {}),
    ("d2-gabor-128x128", "gabor", (128, 128), {}),
    ("d2-mixture-128x128", "mixture", (128, 128), {}),
    ("d2-zone-128x128", "zone", (128, 128), {}),
    ("d2-local-128x128", "local", (128, 128), {}),
    ("d2-zone-256x256", "zone", (256, 256), {}),
    ("d2-mixture-256x256", "mixture", (256, 256), {}),
    ("d2-gabor-256x256", "gabor", (256, 256), {}),
    ("d2-batch3-40x36", "gabor", (3, 40, 36), {"axes": (1, 2)}),
    ("d2-chunk1-24x31", "mixture", (5, 24, 31), {"axes": (1, 2)}),
    ("d2-chunk2-24x31", "mixture", (5, 24, 31), {"axes": (1, 2)}),
    ("d2-chunk5-24x31", "mixture", (5, 24, 31), {"axes": (1, 2)}),
    ("d2-batch70-16x12", "fringes", (70, 16, 12), {"axes": (1, 2)}),
    ("d2-batch70-chunk35", "fringes", (70, 16, 12), {"axes": (1, 2)}),
    ("d2-batch70-local", "local", (70, 16, 12), {"axes": (1, 2)}),
    ("d2-batchmid-24x31", "tones", (24, 3, 31), {"axes": (0, 2)}),
    ("d2-twobatch-40x36", "gabor", (2, 40, 3, 36), {"axes": (1, 3)}),
    ("d2-axes-reversed", "local", (24, 31), {"axes": (1, 0)}),
    ("d2-nan-status", "mixture", (3, 24, 31), {"axes": (1, 2), "nan_at": (0,)}),
    ("d2-constant-element", "gabor", (3, 24, 31), {"axes": (1, 2), "constant_at": (1,)}),
    ("d2-inf-nocheck", "mixture", (2, 24, 31), {"axes": (1, 2), "inf_at": (1,)}),
    ("d2-gain-p40", "gabor", (40, 36), {"gain_pow2": 40, "seed_from": "d2-gabor-40x36"}),
    ("d2-gain-1e-30", "gabor", (40, 36), {"gain": 1e-30, "seed_from": "d2-gabor-40x36"}),
    ("d2-gain-1e300", "gabor", (40, 36), {"gain": 1e300, "dtypes": _F64, "seed_from": "d2-gabor-40x36"}),
    ("d2-negative", "gabor", (40, 36), {"gain": -1.0, "seed_from": "d2-gabor-40x36"}),
    ("d2-offset", "gabor", (40, 36), {"offset": 100.0, "seed_from": "d2-gabor-40x36"}),
    ("d2-maximf1", "mixture", (64, 64), {}),
    ("d2-maximf2", "mixture", (64, 64), {}),
    ("d2-tie0", "mixture", (40, 36), {"quantum": 0.0625}),
    ("d2-quantized", "mixture", (40, 36), {"quantum": 0.0625, "seed_from": "d2-tie0"}),
    ("d2-tie2m20", "mixture", (40, 36), {"quantum": 2.0**-12}),
    ("d2-tie1e-3", "edges", (40, 36), {}),
    ("d3-constant-8x7x6", "constant", (8, 7, 6), {}),
    ("d3-noise-8x7x6", "noise", (8, 7, 6), {}),
    ("d3-fringes-8x7x6", "fringes", (8, 7, 6), {}),
    ("d3-blobs-16x13x11", "blobs", (16, 13, 11), {"params": _MANY}),
    ("d3-tones-16x13x11", "tones", (16, 13, 11), {}),
    ("d3-gabor-20x18x16", "gabor", (20, 18, 16), {}),
    ("d3-fringes-20x18x16", "fringes", (20, 18, 16), {}),
    ("d3-edges-20x18x16", "edges", (20, 18, 16), {}),
    ("d3-flat-20x18x16", "flat", (20, 18, 16), {}),
    ("d3-mixture-32x32x32", "mixture", (32, 32, 32), {}),
    ("d3-gabor-32x32x32", "gabor", (32, 32, 32), {}),
    ("d3-zone-32x32x32", "zone", (32, 32, 32), {}),
    ("d3-local-32x32x32", "local", (32, 32, 32), {}),
    ("d3-gabor-64x64x64", "gabor", (64, 64, 64), {"dtypes": _F32}),
    ("d3-mixture-64x64x64", "mixture", (64, 64, 64), {"dtypes": _F32}),
    ("d3-batch3-20x18x16", "gabor", (3, 20, 18, 16), {"axes": (1, 2, 3)}),
    ("d3-chunk1-16x13x11", "mixture", (5, 16, 13, 11), {"axes": (1, 2, 3)}),
    ("d3-chunk2-16x13x11", "mixture", (5, 16, 13, 11), {"axes": (1, 2, 3)}),
    ("d3-chunk5-16x13x11", "mixture", (5, 16, 13, 11), {"axes": (1, 2, 3)}),
    ("d3-batch9-32x32x32", "fringes", (9, 32, 32, 32), {"axes": (1, 2, 3)}),
    ("d3-batchmid-16x13x11", "tones", (16, 2, 13, 11), {"axes": (0, 2, 3)}),
    ("d3-nan-status", "mixture", (3, 16, 13, 11), {"axes": (1, 2, 3), "nan_at": (2,)}),
    ("d3-inf-nocheck", "mixture", (2, 16, 13, 11), {"axes": (1, 2, 3), "inf_at": (0,)}),
    ("d3-gain-m40", "gabor", (20, 18, 16), {"gain_pow2": -40, "seed_from": "d3-gabor-20x18x16"}),
    ("d3-gain-1e30", "gabor", (20, 18, 16), {"gain": 1e30, "seed_from": "d3-gabor-20x18x16"}),
    ("d3-negative", "gabor", (20, 18, 16), {"gain": -1.0, "seed_from": "d3-gabor-20x18x16"}),
    ("d3-maximf1", "mixture", (20, 18, 16), {}),
    ("d3-maximf2", "mixture", (20, 18, 16), {}),
    ("d3-tie0", "mixture", (20, 18, 16), {"quantum": 0.125}),
    ("d1-growth-100-v7", "growth", (100,), {"params": {"frac": (0.4, 0.5)}}),
    ("d1-growth-512-v1", "growth", (512,), {"params": {"frac": (0.4, 0.5)}}),
    ("d1-tones-100-v4", "tones", (100,), {"params": {"base": 3.0, "ratio": 2.5}}),
    ("d2-growth-40x36-v4", "growth", (40, 36), {}),
    ("d2-growth-64x64-v3", "growth", (64, 64), {"params": {"frac": (0.4, 0.5)}}),
    ("d2-finepart-64x64", "local", (64, 64), {"params": {**_FINE, "coarse_period": 0.2}}),
    ("d2-finepart-128x128-v7", "local", (128, 128), {"params": {**_FINE, "coarse_period": 0.1}}),
    ("d3-growth-32x32x32-v3", "growth", (32, 32, 32), {"params": {"frac": (0.4, 0.5)}}),
    ("d3-finepart-32x32x32-v10", "local", (32, 32, 32), {"params": _FINE}),
    ("d3-gaborone-32x32x32", "gabor", (32, 32, 32),
     {"params": {"count": 1, "p_lo": 3.0, "p_hi": 4.0, "width": (1.0, 1.5)}}),
]
BY_NAME = {r[0]: r for r in RECIPES}


def elements(shape: tuple[int, ...], axes: tuple[int, ...]):
    """(element number, the index of the element in the array, its spatial extents in the order of
    axes), the elements in C order over the batch axes."""
    batch_axes = [i for i in range(len(shape)) if i not in axes]
    n = tuple(shape[a] for a in axes)
    for e, bidx in enumerate(np.ndindex(*[shape[i] for i in batch_axes])):
        idx: list = [slice(None)] * len(shape)
        for i, b in zip(batch_axes, bidx):
            idx[i] = b
        yield e, tuple(idx), n


def build(name: str, dtype: str) -> np.ndarray:
    """The input of the manifest case (recipe name, dtype)."""
    _, family, shape, opt = BY_NAME[name]
    if dtype not in opt.get("dtypes", ("float32", "float64")):
        raise ValueError(f"{name} has no {dtype} case")
    axes = opt.get("axes") or tuple(range(len(shape)))
    seed = name_seed(opt.get("seed_from") or name)
    make = FAMILIES[family]
    x = np.empty(shape, dtype=np.float64)
    for e, idx, n in elements(shape, axes):
        f = np.asarray(make(n, Draws(seed, e), **opt.get("params", {})), dtype=np.float64)
        assert f.shape == n, (name, f.shape, n)
        x[idx] = f.transpose([axes.index(a) for a in sorted(axes)])
    if opt.get("quantum") is not None:
        x = trunc_floor(x / opt["quantum"]) * opt["quantum"]
    if opt.get("gain_pow2") is not None:
        x = np.ldexp(x, opt["gain_pow2"])
    if opt.get("gain") is not None:
        x = x * opt["gain"]
    if opt.get("offset") is not None:
        x = x + opt["offset"]
    x = x.astype(dtype)
    for e, idx, _ in elements(shape, axes):
        if e in opt.get("constant_at", ()):
            x[idx] = np.asarray(0.75, dtype=dtype)
        view = x[idx]
        pos = np.unravel_index((view.size * 2) // 5, view.shape)
        if e in opt.get("nan_at", ()):
            view[pos] = np.nan
        if e in opt.get("inf_at", ()):
            view[pos] = np.inf
    return x


def sha256(a: np.ndarray) -> str:
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()
