# SPDX-License-Identifier: Apache-2.0
"""The adjoint (eda_backward through eda.differentiable_decompose) on the GPU.

- The frozen forward (eda_replay) at the call's own input gives the call's rows bit for bit.
- Against the FP64 reference (oracle/adjoint_ref.py: the frozen sift map restated in PyTorch
  float64 from the oracle's tape, differentiated by autograd): the loss sum_k w_k ||row_k -
  target_k||^2 (w_k ~ U(0.5, 2), targets ~ N(0, 1)); the GPU's float64 gradient within 1e-8
  relative L2 of the reference, its float32 gradient within 1e-3 (float32 rows round at 1e-7 and a
  near-tie of a max / min can route a voxel's gradient to the other candidate), with and without
  the orientation's derivative.
- Against FP64 central differences (eps = 1e-3) of the GPU's frozen forward at 64 pixels of the
  differentiation workload (32 x 256 x 256 images): pixels where the central differences at eps and
  eps / 2 disagree beyond O(eps^2) (a max / min tie crossed inside +-eps) are skipped; the
  adjoint's derivative within 1e-4 relative of the central difference at the others.
- The float32 gradient at full size within 1e-3 relative L2 of the float64 one, same decisions.
- autograd: torch.autograd.grad through eda.differentiable_decompose equals eda_backward up to the
  order of the gather transpose's atomic additions (1e-6 relative L2).
"""

from __future__ import annotations

import numpy as np
import pytest

import adjoint_ref
import eda
import oracle
import workloads

pytestmark = pytest.mark.gpu


def torch_cuda():
    import torch

    assert torch.cuda.is_available()
    return torch


def rel_l2(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


def loss_grad_rows(torch, rows, seed):
    """d loss / d rows for the workload's loss, per element (the same w, targets for every element)."""
    P = rows.shape[1]
    w, tgt = adjoint_ref.loss_weights(P, seed, tuple(rows.shape[2:]))
    w = w.to(rows.device, rows.dtype).reshape((1, P) + (1,) * (rows.dim() - 2))
    tgt = tgt.to(rows.device, rows.dtype)[None]
    return 2.0 * w * (rows - tgt), w, tgt


@pytest.mark.parametrize("dtype", ["float32", "float64"])
@pytest.mark.parametrize("shape", [(3, 1024), (3, 48, 56), (2, 20, 22, 18)])
def test_replay_gives_the_rows_bitwise(dtype, shape):
    torch = torch_cuda()
    print(f"adjoint: replay {dtype} {shape}", flush=True)
    rng = np.random.default_rng(3)
    if len(shape) == 3:
        x = workloads.rings_stripes_blobs(shape[0], 56)[:, : shape[1], : shape[2]]
    else:
        x = rng.standard_normal(shape)
        x = x + np.cumsum(rng.standard_normal(shape), axis=-1) * 0.1
    xt = torch.from_numpy(np.ascontiguousarray(x).astype(dtype)).to("cuda")
    td = eda.TapedDecomposition(xt)
    rows = td.replay(xt)
    assert torch.equal(rows, td.rows)
    y = eda.decompose(xt, axes=tuple(range(1, len(shape))))
    assert torch.equal(y, td.rows)


@pytest.mark.parametrize("orientation_grad", [True, False])
@pytest.mark.parametrize("dtype,tol", [("float64", 1e-8), ("float32", 1e-3)])
def test_gradient_matches_the_fp64_reference(dtype, tol, orientation_grad):
    torch = torch_cuda()
    print(f"adjoint: vs reference {dtype} orientation_grad={orientation_grad}", flush=True)
    x = workloads.rings_stripes_blobs(2, 48).astype(np.float64)
    x[1] = x[1] * 3.0 + 0.5
    xt = torch.from_numpy(x.astype(dtype)).to("cuda")
    td = eda.TapedDecomposition(xt)
    for b in range(2):
        tape = oracle.decompose_tape(x[b])
        k = int(td.info["n_imfs"][b])
        assert k == tape["n_imfs"]
        assert td.info["windows"][b, :k].tolist() == [m["w"] for m in tape["modes"]]
        loss, gref, rows_ref = adjoint_ref.reference_gradient(x[b], tape, seed=5 + b,
                                                              orientation_grad=orientation_grad)
        rb = td.rows[b:b + 1, : k + 1]
        g, _, _ = loss_grad_rows(torch, rb, 5 + b)
        full = torch.zeros_like(td.rows[b:b + 1])
        full[:, : k + 1] = g
        grows = torch.zeros_like(td.rows)
        grows[b:b + 1] = full
        gx = td.backward(grows, orientation_grad=orientation_grad)[b].cpu().numpy()
        err = rel_l2(gx, gref)
        print(f"  element {b}: rel L2 {err:.3e}", flush=True)
        assert err <= tol, (b, err)


def test_gradient_matches_central_differences_on_the_differentiation_workload():
    torch = torch_cuda()
    print("adjoint: central differences, 32 x 256 x 256 float64", flush=True)
    x = workloads.rings_stripes_blobs(32, 256).astype(np.float64)
    xt = torch.from_numpy(x).to("cuda")
    td = eda.TapedDecomposition(xt)
    seed = 11
    g, w, tgt = loss_grad_rows(torch, td.rows, seed)
    mask = torch.zeros_like(td.rows)
    for b in range(32):
        mask[b, : int(td.info["n_imfs"][b]) + 1] = 1.0
    g = g * mask
    gx = td.backward(g).cpu().numpy()

    def losses(xx):
        rows = td.replay(xx)
        return ((w * (rows - tgt) ** 2) * mask).sum(dim=tuple(range(1, rows.dim()))).cpu().numpy()

    rng = np.random.default_rng(19)
    eps = 1e-3
    checked, worst = 0, 0.0
    for rnd in range(6):
        pix = rng.integers(0, 256, (32, 2))
        dx = torch.zeros_like(xt)
        for b in range(32):
            dx[b, pix[b, 0], pix[b, 1]] = eps
        d1 = (losses(xt + dx) - losses(xt - dx)) / (2 * eps)
        d2 = (losses(xt + 0.5 * dx) - losses(xt - 0.5 * dx)) / eps
        for b in range(32):
            # a smooth loss gives central differences at eps and eps / 2 that agree to O(eps^2); a
            # max / min tie crossed inside +-eps makes them differ: no derivative to compare there
            if abs(d1[b] - d2[b]) > 1e-5 * (abs(d1[b]) + 1e-2):
                continue
            ad = gx[b, pix[b, 0], pix[b, 1]]
            rel = abs(ad - d1[b]) / max(abs(d1[b]), 1e-2)
            worst = max(worst, rel)
            checked += 1
            assert rel <= 1e-4, (b, pix[b], ad, d1[b])
        if checked >= 64:
            break
    print(f"  {checked} pixels, worst relative error {worst:.2e}", flush=True)
    assert checked >= 64


def test_float32_gradient_tracks_float64_at_full_size():
    torch = torch_cuda()
    print("adjoint: float32 vs float64 gradients, 32 x 256 x 256", flush=True)
    x = workloads.rings_stripes_blobs(32, 256)
    grads = []
    for dt in (torch.float32, torch.float64):
        xt = torch.from_numpy(x).to("cuda", dt)
        td = eda.TapedDecomposition(xt)
        g, _, _ = loss_grad_rows(torch, td.rows, 7)
        grads.append((td.info["n_imfs"].copy(), td.backward(g).double().cpu().numpy()))
    assert np.array_equal(grads[0][0], grads[1][0])
    err = rel_l2(grads[0][1], grads[1][1])
    print(f"  rel L2 {err:.3e}", flush=True)
    assert err <= 1e-3


def test_autograd_function_equals_the_backward():
    torch = torch_cuda()
    x = torch.from_numpy(workloads.rings_stripes_blobs(3, 64)).to("cuda").requires_grad_(True)
    rows = eda.differentiable_decompose(x)
    w = torch.linspace(0.5, 2.0, rows.shape[1], device="cuda").reshape(1, -1, 1, 1)
    loss = (w * rows * rows).sum()
    (gx,) = torch.autograd.grad(loss, x)
    td = eda.TapedDecomposition(x.detach())
    ref = td.backward(2.0 * w * td.rows)
    assert rel_l2(gx.cpu().numpy(), ref.cpu().numpy()) <= 1e-6
