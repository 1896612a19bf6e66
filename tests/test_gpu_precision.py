# SPDX-License-Identifier: Apache-2.0
"""The precision form (dct="bf16x3"): every DCT / Riesz contraction of the orientation as split-BF16
products on the matrix cores (three bf16 products per term, float32 accumulation), on the GPU.

Reported per mode against float64 (the GPU's float64 decomposition, itself within 1e-11 of the FP64
oracle) on the precision workload (4 x 256^3 float32 plane-wave volumes): the decisions (n_imfs,
windows, sifts) equal float64's, and each row's relative L2 error stays within 1e-3 (measured
4.4e-4 at worst, mode 1; the exact float32 transforms give 2.9e-5 there). On 2-D images the
orientation's coherence and direction stay within 1e-3 of the exact float32 transforms', and the
rows within 1e-4 relative L2 of float64. dct="bf16x3" on float64 input runs the exact float64
transforms (the bits of dct="exact")."""

from __future__ import annotations

import numpy as np
import pytest

import eda
import workloads

pytestmark = pytest.mark.gpu


def torch_cuda():
    import torch

    assert torch.cuda.is_available()
    return torch


def rel(torch, a, b):
    return float(torch.linalg.vector_norm((a - b).double()) / torch.linalg.vector_norm(b.double()).clamp_min(1e-300))


def per_mode_errors(torch, y, info, y64):
    errs = []
    for k in range(y.shape[1]):
        es = [rel(torch, y[b, k], y64[b, k]) for b in range(y.shape[0]) if k <= int(info.n_imfs[b])]
        errs.append(max(es) if es else 0.0)
    return errs


def test_precision_volumes_per_mode_error_against_float64():
    torch = torch_cuda()
    print("precision: 4 x 256^3 float32, bf16x3 vs float64 per mode", flush=True)
    x = torch.from_numpy(workloads.plane_wave_volumes(4, 256)).to("cuda")
    y64, i64 = eda.decompose(x.double(), axes=(1, 2, 3), return_info=True)
    yb, ib = eda.decompose(x, axes=(1, 2, 3), return_info=True, dct="bf16x3")
    assert np.array_equal(ib.n_imfs, i64.n_imfs)
    assert np.array_equal(ib.windows, i64.windows) and np.array_equal(ib.sifts, i64.sifts)
    errs = per_mode_errors(torch, yb, ib, y64)
    print("  per-mode max rel L2: " + ", ".join(f"{e:.2e}" for e in errs), flush=True)
    assert max(errs) <= 1e-3


def test_bf16x3_orientation_tracks_the_exact_transforms():
    torch = torch_cuda()
    x = torch.from_numpy(workloads.rings_stripes_blobs(4, 256)).to("cuda")
    ye, ie, oe = eda.decompose(x, axes=(1, 2), return_info=True, return_orientation=True)
    yb, ib, ob = eda.decompose(x, axes=(1, 2), return_info=True, return_orientation=True, dct="bf16x3")
    assert np.array_equal(ie.n_imfs, ib.n_imfs) and np.array_equal(ie.windows, ib.windows)
    c_err = float((oe.coherence - ob.coherence).abs().max())
    strong = oe.coherence > 0.2
    dots = (oe.direction * ob.direction).sum(dim=2).abs()
    n_err = float((1.0 - dots)[strong].max())
    print(f"precision: 2-D orientation |dc| {c_err:.2e}, 1 - |n.n'| {n_err:.2e}", flush=True)
    assert c_err <= 1e-3 and n_err <= 1e-3
    y64 = eda.decompose(x.double(), axes=(1, 2))
    errs = per_mode_errors(torch, yb, ib, y64)
    assert max(errs) <= 1e-4, errs


def test_bf16x3_on_float64_input_is_the_exact_path():
    torch = torch_cuda()
    x = torch.from_numpy(workloads.rings_stripes_blobs(2, 64).astype(np.float64)).to("cuda")
    assert torch.equal(eda.decompose(x, axes=(1, 2), dct="bf16x3"), eda.decompose(x, axes=(1, 2)))


def test_bf16x3_adjoint_replays_its_forward():
    torch = torch_cuda()
    x = torch.from_numpy(workloads.rings_stripes_blobs(2, 64)).to("cuda")
    td = eda.TapedDecomposition(x, dct="bf16x3")
    assert torch.equal(td.replay(x), td.rows)
    assert torch.equal(td.rows, eda.decompose(x, axes=(1, 2), dct="bf16x3"))
