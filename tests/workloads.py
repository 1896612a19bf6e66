# SPDX-License-Identifier: Apache-2.0
"""Synthetic inputs of the benchmark and parity workloads (test and benchmark infrastructure).

- tones_2048: float32 length 2048, x[t] = sin(2 pi t/8) + 2 sin(2 pi t/64) + 3 sin(2 pi t/512) +
  N(0, 0.01^2) noise from numpy.random.default_rng(0).
- rings_stripes_blobs(B, n): B float32 n x n images; image i from default_rng(i), drawn in order:
  the ring centre (cy, cx) ~ U(0, n)^2, the stripe angle theta ~ U(0, pi), two blobs (centre ~
  U(0, n)^2, amplitude ~ U(1, 3)), then N(0, 0.05^2) noise; the field is cos(2 pi r / 6) +
  cos(2 pi (x' + 8 sin(2 pi y' / 128)) / 24) with (x', y') the coordinates rotated by theta, plus
  the blobs amp exp(-|p - c|^2 / (2 80^2)) and the noise, computed in float64, stored as float32.
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
