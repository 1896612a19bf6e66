# SPDX-License-Identifier: Apache-2.0
"""The precision workload (4 x 256^3 float32 plane-wave volumes) and the differentiation workload's
forward pass (32 x 256^2 float32 images) on the GPU (device input), at full size by properties and
on samples against the FP64 oracle.

Full size: no element fails; every row is finite; the rows reconstruct the input within
64 * 2^-24 * max sum |rows| (the float32 sum of P rows in any order); each element decomposed alone
gives its batched rows bit for bit (batch invariance, reference driver.cpp:754-923 runs elements
independently); a call limited to one element per chunk gives the same bits (chunking invariance).
Against the FP64 oracle (float32 on the GPU): decisions (n_imfs, windows, sifts) exact, each row
within 1e-5 relative L2 of the oracle's (the float32 path's tolerance). For the volumes the oracle
check runs on a 64^3 crop of volume 0 around its wave envelope (the whole volume takes the oracle
minutes); for the images on images 0 and 17 of the batch."""

from __future__ import annotations

import numpy as np
import pytest

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


def full_size_properties(torch, x, axes):
    """Decompose the device batch x; check the properties above; return (rows, info)."""
    xt = torch.from_numpy(x).to("cuda")
    y, info = eda.decompose(xt, axes=axes, return_info=True)
    assert (info.error == 0).all(), info.error
    assert bool(torch.isfinite(y).all())
    B = x.shape[0]
    red = tuple(range(1, y.dim() - 1))
    ys = y.sum(dim=1)
    S = y.abs().sum(dim=1).amax(dim=red).cpu().numpy()
    dev = (ys - xt).abs().amax(dim=red).cpu().numpy()
    for b in range(B):
        assert dev[b] <= 64 * 2.0**-24 * S[b], (b, dev[b], S[b])
    for b in range(B):
        yb, ib = eda.decompose(xt[b], return_info=True)
        k = int(ib.n_imfs)
        assert k == int(info.n_imfs[b])
        assert ib.windows.tolist() == info.windows[b, :k].tolist()
        assert torch.equal(yb, y[b, : k + 1]), b
    # one element per chunk through the native interface (the reference's max_chunk)
    from eda import _core_hip

    owner, ptr, raw = _core_hip.decompose_device(
        xt.data_ptr(), 0, "float32", "float32", (B,), (int(np.prod(x.shape[1:])),), x.shape[1:],
        tuple(int(np.prod(x.shape[a + 2:])) for a in range(len(axes))),
        [int(torch.cuda.current_stream().cuda_stream)], -1, None, 3, True, False, 1)
    assert raw["chunk"] == 1
    assert np.array_equal(raw["windows"], info.windows) and np.array_equal(raw["n_imfs"], info.n_imfs)
    assert torch.equal(owner, y)
    return y, info


def against_the_oracle(y, info, x, dim):
    yo, io = oracle.decompose(x.astype(np.float64), dim=dim)
    k = int(io["n_imfs"][0])
    assert int(np.asarray(info.n_imfs).reshape(-1)[0]) == k
    assert np.asarray(info.windows).reshape(1, -1)[0, :k].tolist() == io["windows"][0].tolist()
    assert np.asarray(info.sifts).reshape(1, -1)[0, :k].tolist() == io["sifts"][0].tolist()
    yc = np.asarray(y, np.float64)
    for j in range(k + 1):
        assert rel_l2(yc[j], yo[0, j]) <= 1e-5, (j, rel_l2(yc[j], yo[0, j]))


def test_precision_volumes_full_size():
    torch = torch_cuda()
    print("configs: 4 x 256^3 float32 plane-wave volumes", flush=True)
    x = workloads.plane_wave_volumes(4, 256)
    y, info = full_size_properties(torch, x, (1, 2, 3))
    assert (info.n_imfs >= 2).all(), info.n_imfs


def test_precision_crop_against_the_fp64_oracle():
    torch = torch_cuda()
    print("configs: 64^3 crop of plane-wave volume 0, float32 GPU vs FP64 oracle", flush=True)
    c = workloads.plane_wave_crop(0)
    y, info = eda.decompose(torch.from_numpy(c).to("cuda"), return_info=True)
    assert int(info.error) == 0
    against_the_oracle(y.cpu().numpy(), info, c[None], 3)


def test_differentiation_batch_forward():
    torch = torch_cuda()
    print("configs: 32 x 256^2 float32 images", flush=True)
    x = workloads.rings_stripes_blobs(32, 256)
    y, info = full_size_properties(torch, x, (1, 2))
    yc = y.cpu().numpy()
    for b in (0, 17):
        k = int(info.n_imfs[b])
        sub = type("I", (), {"n_imfs": info.n_imfs[b:b + 1], "windows": info.windows[b:b + 1],
                             "sifts": info.sifts[b:b + 1]})
        against_the_oracle(yc[b, : k + 1], sub, x[b:b + 1], 2)
