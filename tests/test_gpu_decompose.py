# SPDX-License-Identifier: Apache-2.0
"""eda.decompose on the GPU against the reference's golden corpus and the FP64 oracle (GPU).

- Golden corpus (FP64): decisions (n_imfs, windows, sifts) exact; rows within the corpus's bound
  A * 1000 * 2^-53 * max|x| (reference tests/python/test_golden.py:119-137).
- float32 on the GPU against the FP64 oracle: decisions exact, modes and residue within 1e-5
  relative L2 (the tolerance of the float32 path).
- float64 on the GPU against the FP64 oracle: decisions exact, rows within 1e-11 relative L2 (the
  device and host libm's log / exp and the DCT's summation order are the only differences).
- Properties at the benchmark's full size (64 x 512 x 512 float32): the rows reconstruct the input,
  every element's decisions equal those of the same element decomposed alone (batch invariance),
  and a sample of elements matches the FP64 oracle.
"""

from __future__ import annotations

from pathlib import Path

import numpy as np
import pytest

import eda
import oracle
import workloads

pytestmark = pytest.mark.gpu
GOLDEN = Path(__file__).resolve().parent / "golden" / "corpus"
CASES = sorted(p.stem for p in GOLDEN.glob("*.npz"))


def torch_cuda():
    import torch

    assert torch.cuda.is_available()
    return torch


def rel_l2(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


@pytest.mark.parametrize("name", CASES)
def test_golden_corpus_on_the_gpu(name):
    with np.load(GOLDEN / f"{name}.npz", allow_pickle=False) as z:
        x, modes, windows, sifts, sens = z["x"], z["modes"], z["windows"], z["sifts"], float(z["sensitivity"])
    y, info = eda.decompose(x, return_info=True)
    assert info.error == 0 and info.dtype == "float64"
    assert y.shape == modes.shape
    assert int(info.n_imfs) == modes.shape[0] - 1
    assert info.windows.tolist() == windows.tolist()
    assert info.sifts.tolist() == sifts.tolist()
    bound = sens / 1e-12 * 1000 * 2.0**-53 * float(np.abs(x).max())
    assert float(np.abs(y - modes).max()) <= bound


def test_tones_float32_device_against_the_fp64_oracle():
    torch = torch_cuda()
    x = workloads.tones_2048(np.float32)
    y, info = eda.decompose(torch.from_numpy(x).to("cuda"), return_info=True)
    yo, io = oracle.decompose(x.astype(np.float64))
    assert isinstance(y, torch.Tensor) and y.device.type == "cuda"
    assert int(info.n_imfs) == int(io["n_imfs"][0])
    assert info.windows.tolist() == io["windows"][0].tolist()
    assert rel_l2(y.cpu().numpy(), yo[0]) <= 1e-5
    for k in range(yo.shape[1]):
        assert rel_l2(y[k].cpu().numpy(), yo[0, k]) <= 1e-5, k


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_2d_batch_against_the_oracle(dtype):
    x = workloads.rings_stripes_blobs(4, 96).astype(dtype)
    y, info = eda.decompose(x, axes=(1, 2), return_info=True)
    yo, io = oracle.decompose(x.astype(np.float64), dim=2)
    assert info.n_imfs.tolist() == io["n_imfs"].tolist()
    assert info.windows.tolist() == io["windows"].tolist()
    tol = 1e-5 if dtype == np.float32 else 1e-11
    for b in range(4):
        assert rel_l2(y[b], yo[b]) <= tol, b


def test_3d_float64_against_the_oracle():
    rng = np.random.default_rng(11)
    g = np.linspace(0, 4 * np.pi, 20)
    zz, yy, xx = np.meshgrid(g, g, g, indexing="ij")
    x = np.sin(2.1 * xx + 0.7 * yy) + 0.5 * np.cos(0.8 * zz - 1.3 * xx) + 0.05 * rng.standard_normal(xx.shape)
    y, info = eda.decompose(x, return_info=True)
    yo, io = oracle.decompose(x)
    assert int(info.n_imfs) == int(io["n_imfs"][0])
    assert info.windows.tolist() == io["windows"][0].tolist()
    assert rel_l2(y, yo[0]) <= 1e-11


def test_1d_float64_device_is_the_oracle_to_libm_rounding():
    torch = torch_cuda()
    x = np.stack([workloads.tones_2048(np.float64), np.load(GOLDEN / "chirp_2048.npz")["x"]])
    y, info = eda.decompose(torch.from_numpy(x).to("cuda"), axes=1, return_info=True)
    yo, io = oracle.decompose(x, dim=1)
    assert info.n_imfs.tolist() == io["n_imfs"].tolist()
    assert info.windows.tolist() == io["windows"].tolist()
    assert info.sifts.tolist() == io["sifts"].tolist()
    yc = y.cpu().numpy()
    for b in range(2):
        assert rel_l2(yc[b], yo[b]) <= 1e-12


def test_edge_cases():
    # two samples, a constant, NaN with check_finite, mixed batch with errors="status"
    y, info = eda.decompose(np.array([1.0, 2.0]), return_info=True)
    assert int(info.n_imfs) == 0 and np.array_equal(y[0], [1.0, 2.0])
    y, info = eda.decompose(np.full(100, 3.25), return_info=True)
    assert int(info.n_imfs) == 0 and np.array_equal(y[0], np.full(100, 3.25))
    x = np.stack([np.sin(np.arange(300.0) / 3), np.sin(np.arange(300.0) / 5)])
    x[0, 7] = np.nan
    with pytest.raises(eda.DecompositionError):
        eda.decompose(x, axes=1)
    y, info = eda.decompose(x, axes=1, errors="status", return_info=True)
    assert info.error.tolist()[0] == 1 and np.isnan(y[0]).all()
    yo, io = oracle.decompose(x[1:], dim=1)
    assert int(info.n_imfs[1]) == int(io["n_imfs"][0])
    assert rel_l2(y[1, : io["n_imfs"][0] + 1], yo[0]) <= 1e-12
    # an empty batch
    y, info = eda.decompose(np.zeros((0, 64)), axes=1, return_info=True)
    assert y.shape == (0, 1, 64)


def test_layout_and_batch_invariance():
    x = workloads.rings_stripes_blobs(3, 64)
    xt = np.ascontiguousarray(np.moveaxis(x, 0, 2))  # (64, 64, 3): batch last
    y, info = eda.decompose(xt, axes=(0, 1), return_info=True)
    assert y.shape[:2] == (y.shape[0], 64) and y.shape[-1] == 3
    for b in range(3):
        yb, ib = eda.decompose(x[b], return_info=True)
        assert int(ib.n_imfs) == int(info.n_imfs[b])
        assert np.array_equal(yb[: int(ib.n_imfs) + 1], y[: int(ib.n_imfs) + 1, :, :, b])


def test_reference_core_interface_matches_decompose():
    """eda._core_hip serves the reference's native-module calls (src/eda/_api.py:310-320 calls them
    with max_chunk 0): the same rows and info as eda.decompose, host and device, bitwise."""
    from eda import _core_hip

    torch = torch_cuda()
    x = workloads.rings_stripes_blobs(3, 64, np.float32)
    y, info = eda.decompose(x, axes=(1, 2), return_info=True)
    rows, raw = _core_hip.decompose_host(x, -1, None, 3, True, False, 0)
    assert np.array_equal(rows, y)
    for k in ("n_imfs", "error", "windows", "sifts", "spacing", "energy"):
        assert np.array_equal(raw[k], getattr(info, k)), k
    assert raw["chunk"] == 3
    # a chunk limit changes nothing but the chunk
    rows1, raw1 = _core_hip.decompose_host(x, -1, None, 3, True, False, 1)
    assert np.array_equal(rows1, y) and raw1["chunk"] == 1
    # device input: (3, 64, 64) viewed through a transposed (64, 64, 3) tensor
    xt = torch.from_numpy(np.ascontiguousarray(np.moveaxis(x, 0, 2))).to("cuda")
    owner, ptr, rawd = _core_hip.decompose_device(
        xt.data_ptr(), 0, "float32", "float32", (3,), (1,), (64, 64), (64 * 3, 3),
        [int(torch.cuda.current_stream().cuda_stream)], -1, None, 3, True, False, 0)
    assert ptr == owner.data_ptr() and tuple(owner.shape) == y.shape
    assert np.array_equal(owner.cpu().numpy(), y)
    assert np.array_equal(rawd["windows"], info.windows)
    assert _core_hip.pointer_device(ptr) == 0
    assert _core_hip.device_count() >= 1 and "gfx950" in _core_hip.device_info(0)["gcn_arch"]


def test_full_size_benchmark_batch_properties():
    """64 x 512 x 512 float32 on the device: reconstruction, batch invariance of a sample of
    elements, and one element against the FP64 oracle."""
    torch = torch_cuda()
    x = workloads.rings_stripes_blobs(64, 512)
    xt = torch.from_numpy(x).to("cuda")
    y, info = eda.decompose(xt, axes=(1, 2), return_info=True)
    assert (info.error == 0).all()
    ys = y.sum(dim=1).cpu().numpy()
    S = y.abs().sum(dim=1).amax(dim=(1, 2)).cpu().numpy()
    for b in range(64):
        assert np.abs(ys[b] - x[b]).max() <= 64 * 2.0**-24 * S[b]
    for b in (0, 37):
        yb, ib = eda.decompose(xt[b], return_info=True)
        assert int(ib.n_imfs) == int(info.n_imfs[b])
        assert ib.windows.tolist() == info.windows[b, : int(ib.n_imfs)].tolist()
        assert torch.equal(yb, y[b, : int(ib.n_imfs) + 1])
    yo, io = oracle.decompose(x[5:6].astype(np.float64), dim=2)
    assert int(info.n_imfs[5]) == int(io["n_imfs"][0])
    assert info.windows[5, : int(info.n_imfs[5])].tolist() == io["windows"][0].tolist()
    assert rel_l2(y[5, : int(info.n_imfs[5]) + 1].cpu().numpy(), yo[0]) <= 1e-5
