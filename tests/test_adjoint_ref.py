# SPDX-License-Identifier: Apache-2.0
"""The adjoint's FP64 reference (oracle/adjoint_ref.py) pinned on the CPU.

- Its frozen map, replayed from the oracle's tape, gives the oracle's rows (themselves pinned by the
  reference's golden corpus and parity decisions) to within 1e-13 relative L2, in 1-D, 2-D and 3-D:
  the restatement computes the same function the decomposition does at its decisions.
- The tape's decisions are the oracle's: the windows per mode are the decomposition's, and the
  sifts counted (1 + the accepted ones) are its sift counts.
- Autograd of that map against FP64 central differences (eps = 1e-3, differences at eps and eps / 2
  agreeing, i.e. no max / min tie inside the step) on a 40 x 40 image: within 1e-5 relative.
- The adjoint of each building block against its closed form: the cube's transpose routes the
  gradient to the sliding window's arg-max (numpy), the box transpose is the box matrix's transpose,
  the D-linear interpolation's gradient in its coordinate is the difference of the two samples.
"""

from __future__ import annotations

import numpy as np
import pytest
import torch

import adjoint_ref as ar
import oracle
import workloads


def rel_l2(a, b):
    return float(np.linalg.norm(np.asarray(a) - np.asarray(b)) / max(np.linalg.norm(np.asarray(b)), 1e-300))


@pytest.mark.parametrize("kind", ["1d", "2d", "3d"])
def test_frozen_map_replays_the_oracle(kind):
    if kind == "1d":
        x = workloads.tones_2048(np.float64)[:768]
    elif kind == "2d":
        x = workloads.rings_stripes_blobs(1, 40)[0].astype(np.float64)
    else:
        x = workloads.plane_wave_volumes(1, 20)[0].astype(np.float64)
    tape = oracle.decompose_tape(x)
    y, info = oracle.decompose(x[None], dim=x.ndim)
    assert tape["n_imfs"] == int(info["n_imfs"][0]) >= 1
    assert [m["w"] for m in tape["modes"]] == info["windows"][0].tolist()
    assert [1 + sum(m["accepted"]) for m in tape["modes"]] == info["sifts"][0].tolist()
    rows = ar.frozen_decompose(torch.from_numpy(x), tape).numpy()
    assert rows.shape == tape["rows"].shape
    assert rel_l2(rows, y[0]) <= 1e-13


def test_autograd_of_the_frozen_map_matches_central_differences():
    x = workloads.rings_stripes_blobs(1, 40)[0].astype(np.float64)
    tape = oracle.decompose_tape(x)
    L0, g, rows = ar.reference_gradient(x, tape, seed=3)
    w, tgt = ar.loss_weights(rows.shape[0], 3, x.shape)

    def loss(xx):
        with torch.no_grad():
            r = ar.frozen_decompose(torch.from_numpy(xx), tape)
            return float((w.reshape(-1, 1, 1) * (r - tgt) ** 2).sum())

    rng = np.random.default_rng(2)
    eps, checked = 1e-3, 0
    for _ in range(24):
        i = tuple(rng.integers(0, 40, 2))
        e = np.zeros_like(x)
        e[i] = eps
        d1 = (loss(x + e) - loss(x - e)) / (2 * eps)
        d2 = (loss(x + e / 2) - loss(x - e / 2)) / eps
        if abs(d1 - d2) > 1e-5 * (abs(d1) + 1e-2):
            continue
        assert abs(g[i] - d1) <= 1e-5 * max(abs(d1), 1e-2), (i, g[i], d1)
        checked += 1
        if checked >= 8:
            break
    assert checked >= 8


def test_cube_transpose_routes_to_the_window_argmax():
    rng = np.random.default_rng(0)
    f = torch.from_numpy(rng.standard_normal(37)).requires_grad_(True)
    w = 9
    y = ar.cube(f, "max", w)
    gy = torch.from_numpy(rng.standard_normal(37))
    (g,) = torch.autograd.grad((y * gy).sum(), f)
    # closed form: the sliding max over the clipped window and its arg-max, numpy
    v = f.detach().numpy()
    want = np.zeros(37)
    r = (w - 1) // 2
    for xo in range(37):
        lo, hi = max(0, xo - r), min(36, xo + r)
        want[lo + int(np.argmax(v[lo:hi + 1]))] += gy[xo].item()
    assert np.allclose(g.numpy(), want, rtol=0, atol=1e-15)


def test_box_transpose_is_the_matrix_transpose():
    n, w = 11, 7
    eye = torch.eye(n, dtype=torch.float64)
    M = torch.stack([ar.cube(eye[k], "box", w) for k in range(n)], dim=1)  # column k: image of e_k
    rng = np.random.default_rng(1)
    gy = torch.from_numpy(rng.standard_normal(n))
    f = torch.zeros(n, dtype=torch.float64, requires_grad=True)
    (g,) = torch.autograd.grad((ar.cube(f, "box", w) * gy).sum(), f)
    assert np.allclose(g.numpy(), (M.T @ gy).numpy(), rtol=0, atol=1e-15)


def test_interpolation_gradient_in_the_coordinate():
    C = torch.arange(12.0, dtype=torch.float64).reshape(3, 4) ** 2
    y = torch.tensor([[1.25], [2.5]], dtype=torch.float64, requires_grad=True).reshape(2, 1, 1)
    v = ar.interp(C, y)
    (g,) = torch.autograd.grad(v.sum(), y)
    # bilinear in (1.25, 2.5): d/dy0 = lerp along axis 1 of the row difference, d/dy1 likewise
    c = C.numpy()
    d0 = (1 - 0.5) * (c[2, 2] - c[1, 2]) + 0.5 * (c[2, 3] - c[1, 3])
    d1 = (1 - 0.25) * (c[1, 3] - c[1, 2]) + 0.25 * (c[2, 3] - c[2, 2])
    assert np.allclose(g.reshape(2).numpy(), [d0, d1], rtol=0, atol=1e-12)
