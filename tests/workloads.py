# SPDX-License-Identifier: Apache-2.0
"""Synthetic inputs of the benchmark and parity workloads (test and benchmark infrastructure).

- tones_2048: float32 length 2048, x[t] = sin(2 pi t/8) + 2 sin(2 pi t/64) + 3 sin(2 pi t/512) +
  N(0, 0.01^2) noise from numpy.random.default_rng(0).
- rings_stripes_blobs(B, n): B float32 n x n images; image i from default_rng(i), drawn in order:
  the ring centre (cy, cx) ~ U(0, n)^2, the stripe angle theta ~ U(0, pi), two blobs (centre ~
  U(0, n)^2, amplitude ~ U(1, 3)), then N(0, 0.05^2) noise; the field is cos(2 pi r / 6) +
  cos(2 pi (x' + 8 sin(2 pi y' / 128)) / 24) with (x', y') the coordinates rotated by theta, plus
  the blobs amp exp(-|p - c|^2 / (2 80^2)) and the noise, computed in float64, stored as float32.
- plane_wave_volumes(B, n): B float32 n^3 volumes (the precision workload at n = 256); volume i
  from default_rng(i), drawn in order: for the wavelengths 6, 20 and 64 voxels a unit direction
  (three normal draws, normalized) and a phase ~ U(0, 2 pi); the envelope centre ~ U(0, n)^3; the
  trend centre ~ U(0, n)^3 and amplitude ~ U(1, 3); then N(0, 0.02^2) noise. The field is
  exp(-|p - c_e|^2 / (2 60^2)) * sum_k cos(2 pi u_k . p / lambda_k + phi_k) + a_t exp(-|p - c_t|^2 /
  (2 (n / 2)^2)) + noise, computed in float64 per z-slab, stored as float32.
- The differentiation workload is rings_stripes_blobs(32, 256).
- ragged_fields(S, I): the batching workload (S = 256 signals, I = 32 images): S float32 1-D
  signals, signal i from default_rng(i) drawn in order: the length n = round(exp(U(ln 1024,
  ln 262144))), f0 ~ U(0.001, 0.01), f1 ~ U(0.05, 0.2), omega ~ U(0.005, 0.05); the field is the
  linear chirp cos(2 pi (f0 t + (f1 - f0) t^2 / (2 n))) + cos(omega t + 0.8 sin(omega t)) + the
  quadratic trend (t / n)^2; then I float32 images, image j with the side drawn from
  default_rng(10000 + j) uniform in {128, 256, 384, 512, 768, 1024} and built as
  rings_stripes_blobs(1, side, first=10000 + j).
"""

from __future__ import annotations

import numpy as np


def tones_2048(dtype=np.float32) -> np.ndarray:
    t = np.arange(2048, dtype=np.float64)
    x = np.sin(2 * np.pi * t / 8) + 2 * np.sin(2 * np.pi * t / 64) + 3 * np.sin(2 * np.pi * t / 512)
    x = x + np.random.default_rng(0).normal(0.0, 0.01, 2048)
    return x.astype(dtype)


def rings_stripes_blobs(B: int = 64, n: int = 512, dtype=np.float32, first: int = 0) -> np.ndarray:
    out = np.empty((B, n, n), dtype=dtype)
    y, x = np.meshgrid(np.arange(n, dtype=np.float64), np.arange(n, dtype=np.float64), indexing="ij")
    for k in range(B):
        rng = np.random.default_rng(first + k)
        cy, cx = rng.uniform(0, n, 2)
        theta = rng.uniform(0, np.pi)
        blobs = []
        for _ in range(2):
            c = rng.uniform(0, n, 2)
            amp = rng.uniform(1, 3)
            blobs.append((c, amp))
        noise = rng.normal(0.0, 0.05, (n, n))
        r = np.hypot(y - cy, x - cx)
        xr = np.cos(theta) * x + np.sin(theta) * y
        yr = -np.sin(theta) * x + np.cos(theta) * y
        f = np.cos(2 * np.pi * r / 6) + np.cos(2 * np.pi * (xr + 8 * np.sin(2 * np.pi * yr / 128)) / 24)
        for (c, amp) in blobs:
            f = f + amp * np.exp(-((y - c[0]) ** 2 + (x - c[1]) ** 2) / (2 * 80.0**2))
        out[k] = (f + noise).astype(dtype)
    return out


WAVELENGTHS = (6.0, 20.0, 64.0)


def envelope_centre(rng, n):
    return rng.uniform(0, n, 3)


def plane_wave_crop(i: int, n: int = 256, m: int = 64, dtype=np.float32) -> np.ndarray:
    """The m^3 crop of volume i of plane_wave_volumes(., n) centred on its envelope (clipped to the
    volume), where the waves are strongest."""
    v = plane_wave_volumes(1, n, dtype, first=i)[0]
    rng = np.random.default_rng(i)
    for _ in WAVELENGTHS:
        rng.normal(size=3)
        rng.uniform(0.0, 2.0 * np.pi)
    ce = envelope_centre(rng, n)
    o = np.clip(np.round(ce).astype(int) - m // 2, 0, n - m)
    return np.ascontiguousarray(v[o[0]:o[0] + m, o[1]:o[1] + m, o[2]:o[2] + m])


def plane_wave_volumes(B: int = 4, n: int = 256, dtype=np.float32, first: int = 0) -> np.ndarray:
    out = np.empty((B, n, n, n), dtype=dtype)
    ax = np.arange(n, dtype=np.float64)
    yy, xx = np.meshgrid(ax, ax, indexing="ij")
    for k in range(B):
        rng = np.random.default_rng(first + k)
        waves = []
        for lam in WAVELENGTHS:
            u = rng.normal(size=3)
            u = u / np.linalg.norm(u)
            waves.append((u, lam, rng.uniform(0.0, 2.0 * np.pi)))
        ce = envelope_centre(rng, n)
        ct = rng.uniform(0, n, 3)
        at = rng.uniform(1.0, 3.0)
        noise_rng = rng
        st = n / 2.0
        for z in range(n):
            d_e = (z - ce[0]) ** 2 + (yy - ce[1]) ** 2 + (xx - ce[2]) ** 2
            d_t = (z - ct[0]) ** 2 + (yy - ct[1]) ** 2 + (xx - ct[2]) ** 2
            w = np.zeros((n, n))
            for u, lam, phi in waves:
                w += np.cos(2.0 * np.pi * (u[0] * z + u[1] * yy + u[2] * xx) / lam + phi)
            f = np.exp(-d_e / (2.0 * 60.0**2)) * w + at * np.exp(-d_t / (2.0 * st**2))
            out[k, z] = (f + noise_rng.normal(0.0, 0.02, (n, n))).astype(dtype)
    return out


def chirp_signal(i: int, dtype=np.float32, n: int | None = None) -> np.ndarray:
    rng = np.random.default_rng(i)
    drawn = int(round(float(np.exp(rng.uniform(np.log(1024.0), np.log(262144.0))))))
    n = drawn if n is None else n
    f0, f1 = rng.uniform(0.001, 0.01), rng.uniform(0.05, 0.2)
    om = rng.uniform(0.005, 0.05)
    t = np.arange(n, dtype=np.float64)
    x = np.cos(2 * np.pi * (f0 * t + (f1 - f0) * t * t / (2 * n))) + np.cos(om * t + 0.8 * np.sin(om * t))
    return (x + (t / n) ** 2).astype(dtype)


IMAGE_SIDES = (128, 256, 384, 512, 768, 1024)


def ragged_image(j: int, dtype=np.float32) -> np.ndarray:
    side = int(np.random.default_rng(10000 + j).choice(IMAGE_SIDES))
    return rings_stripes_blobs(1, side, dtype, first=10000 + j)[0]


def ragged_fields(S: int = 256, I: int = 32, dtype=np.float32) -> list:
    return [chirp_signal(i, dtype) for i in range(S)] + [ragged_image(j, dtype) for j in range(I)]
