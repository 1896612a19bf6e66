# SPDX-License-Identifier: Apache-2.0
"""Device input of every element type and any batch layout, read in place and cast on the device
(reference cpp/src/field/convert.hpp:17-49, convert.cu; cpp/src/decompose.cpp:25-82;
src/eda/_device.py:20-35).

- The cast table: bool to 0 / 1, integers to the compute dtype with round to nearest even, float16,
  bfloat16 and float32 to float64 exactly. A constant element has no mode, so its one output row is
  its input as cast: every value is compared bit for bit with NumPy's cast (bfloat16: PyTorch's
  exact widening), including values that round (2^24 + 1 in float32, 2^53 + 1 and 2^64 - 1 in
  float64, ties to even).
- A decomposition and a monogenic signal of device input of every type equal those of the host
  input cast by NumPy, bit for bit.
- Batch axes that do not merge into one stride (a transposed or strided batch), read in place,
  give the host bits; so does the reference's native interface with a source dtype.
"""

from __future__ import annotations

import numpy as np
import pytest

import eda

pytestmark = pytest.mark.gpu

TYPES = ["bool", "int8", "int16", "int32", "int64", "uint8", "uint16", "uint32", "uint64", "float16",
         "bfloat16", "float32", "float64"]


def torch_mod():
    import torch

    assert torch.cuda.is_available()
    return torch


def bits(a) -> bytes:
    if hasattr(a, "cpu"):
        a = a.cpu().numpy()
    a = np.ascontiguousarray(a)
    return a.tobytes() + str(a.shape).encode() + str(a.dtype).encode()


def values_of(name: str) -> np.ndarray:
    """Values of the type, including ones that round in the compute dtypes."""
    if name == "bool":
        return np.array([0, 1, 1, 0], dtype=bool)
    if name in ("float16", "bfloat16", "float32", "float64"):
        v = np.array([0.0, -0.0, 1.5, -2.25, 65504.0, 1e-3, 3.0e-5, -7.0], dtype=np.float64)
        return v if name == "bfloat16" else v.astype(name)
    info = np.iinfo(name)
    cand = [0, 1, -1, 100, -100, 2**24 + 1, 2**24 + 3, 2**53 + 1, 2**53 + 3, 2**63 - 1, -(2**63),
            2**64 - 1, 2**62 + 2**9, info.max, info.min]
    return np.array(sorted({c for c in cand if info.min <= c <= info.max}), dtype=name)


def device_tensor(torch, name: str, host: np.ndarray):
    if name == "bfloat16":
        return torch.from_numpy(host.astype(np.float32)).to(torch.bfloat16).cuda()
    return torch.from_numpy(np.ascontiguousarray(host)).cuda()


def host_cast(torch, name: str, host: np.ndarray, target: str) -> np.ndarray:
    if name == "bfloat16":
        return torch.from_numpy(host.astype(np.float32)).to(torch.bfloat16).float().numpy().astype(target)
    return host.astype(target)


@pytest.mark.parametrize("target", ["float32", "float64"])
@pytest.mark.parametrize("name", TYPES)
def test_the_cast_table(name, target):
    """Each value as a constant element (8 samples): the residue row is the cast input."""
    torch = torch_mod()
    print("device input: cast", name, target, flush=True)
    v = values_of(name)
    host = np.repeat(v[:, None], 8, axis=1)  # (values, 8): one constant element per value
    xd = device_tensor(torch, name, host)
    y, info = eda.decompose(xd, axes=1, dtype=target, return_info=True)
    assert not info.n_imfs.any() and not info.error.any()
    got = y[:, 0].cpu().numpy()
    want = host_cast(torch, name, host, target)
    assert got.dtype == np.dtype(target)
    assert bits(got) == bits(want), (name, target, got[:, 0], want[:, 0])


def waves(shape, seed):
    rng = np.random.default_rng(seed)
    g = np.meshgrid(*[np.arange(n, dtype=np.float64) for n in shape], indexing="ij")
    out = np.zeros(shape)
    for period in (5.0, 13.0):
        k = rng.normal(size=len(shape))
        k = k / np.linalg.norm(k) * 2 * np.pi / period
        out += np.cos(sum(ki * gi for ki, gi in zip(k, g)) + rng.uniform(0, 6))
    return out


@pytest.mark.parametrize("name", TYPES)
def test_decomposition_and_monogenic_of_every_type(name):
    torch = torch_mod()
    print("device input: fields", name, flush=True)
    base = np.stack([waves((40, 36), 3 + b) for b in range(3)])
    if name == "bool":
        host = base > 0
    elif name in ("float16", "float32", "float64"):
        host = base.astype(name)
    elif name == "bfloat16":
        host = base
    else:
        info = np.iinfo(name)
        scale = min(1000.0, float(info.max) / 4)
        host = np.clip(np.round(base * scale + (0 if info.min < 0 else 2 * scale)), info.min, info.max).astype(name)
    xd = device_tensor(torch, name, host)
    for target in ("float32", "float64"):
        ref = host_cast(torch, name, host, target)
        y_h, i_h = eda.decompose(ref, axes=(1, 2), return_info=True, errors="status")
        y_d, i_d = eda.decompose(xd, axes=(1, 2), dtype=target, return_info=True, errors="status")
        assert bits(i_d.n_imfs) == bits(i_h.n_imfs) and bits(i_d.windows) == bits(i_h.windows)
        assert bits(y_d) == bits(y_h)
        m_h = eda.monogenic(ref, axes=(1, 2))
        m_d = eda.monogenic(xd, axes=(1, 2), dtype=target)
        for field in ("amplitude", "phase", "frequency", "direction", "coherence"):
            assert bits(getattr(m_d, field)) == bits(getattr(m_h, field)), field


def test_batch_axes_that_do_not_merge():
    """A transposed batch (strides (S, 3 S)) and a strided one, read in place."""
    torch = torch_mod()
    x = np.stack([np.stack([waves((24, 31), 10 * i + j) for j in range(3)]) for i in range(2)])  # (2, 3, 24, 31)
    xd = torch.from_numpy(x).cuda()
    t = xd.permute(1, 0, 2, 3)  # batch (3, 2) with strides (S, 3 S)
    y_d, i_d = eda.decompose(t, axes=(2, 3), return_info=True)
    y_h, i_h = eda.decompose(np.ascontiguousarray(x.transpose(1, 0, 2, 3)), axes=(2, 3), return_info=True)
    assert bits(y_d) == bits(y_h) and bits(i_d.windows) == bits(i_h.windows)
    wide = torch.from_numpy(np.repeat(x, 2, axis=1)).cuda()[:, ::2]  # batch stride 2 S along axis 1
    y_w = eda.decompose(wide, axes=(2, 3))
    assert bits(y_w) == bits(eda.decompose(x, axes=(2, 3)))
    m_d = eda.monogenic(t, axes=(2, 3))
    m_h = eda.monogenic(np.ascontiguousarray(x.transpose(1, 0, 2, 3)), axes=(2, 3))
    assert bits(m_d.phase) == bits(m_h.phase) and bits(m_d.direction) == bits(m_h.direction)


def test_the_native_interface_casts_device_input():
    torch = torch_mod()
    from eda import _core_hip

    host = np.round(np.stack([waves((20, 18), 30 + b) for b in range(4)]) * 500).astype(np.int16)
    xd = torch.from_numpy(host).cuda()
    ref_rows, ref_info = _core_hip.decompose_host(host.astype(np.float32), -1, None, 3, True, False, 0)
    owner, ptr, info = _core_hip.decompose_device(
        xd.data_ptr(), 0, "int16", "float32", (4,), (360,), (20, 18), (18, 1), (), -1, None, 3, True, False, 0)
    assert bits(owner) == bits(ref_rows) and bits(info["windows"]) == bits(ref_info["windows"])
    # batch axes (2, 2) that do not merge: element (i, j) at i * 360 + j * 720
    t = xd.reshape(2, 2, 20, 18).permute(1, 0, 2, 3)
    owner, ptr, info = _core_hip.decompose_device(
        t.data_ptr(), 0, "int16", "float64", (2, 2), (360, 720), (20, 18), (18, 1), (), -1, None, 3, True, False, 0)
    ref, _ = _core_hip.decompose_host(np.ascontiguousarray(host.reshape(2, 2, 20, 18).transpose(1, 0, 2, 3))
                                      .reshape(4, 20, 18).astype(np.float64), -1, None, 3, True, False, 0)
    assert bits(owner) == bits(ref)


def test_input_still_being_written_on_the_null_stream():
    # the input's last writes are queued on the default (null) stream when the call starts: the
    # call must read it after them, as eda.decompose / eda.monogenic promise for the current stream
    import torch

    rng = np.random.default_rng(11)
    host = rng.standard_normal((16, 256, 256)).astype(np.float32)
    x = torch.from_numpy(host).to("cuda")
    torch.cuda.synchronize()
    want = eda.decompose(x, axes=(1, 2))
    want_m = eda.monogenic(x, axes=(1, 2))
    for _ in range(2):
        y = x.clone()
        for _ in range(40):
            y = y * 1.0 + 0.0  # the same values, a fresh buffer written by queued kernels
        assert torch.equal(eda.decompose(y, axes=(1, 2)), want)
        z = x.clone()
        for _ in range(40):
            z = z * 1.0 + 0.0
        got_m = eda.monogenic(z, axes=(1, 2))
        for key in ("amplitude", "phase", "frequency", "direction", "coherence"):
            assert torch.equal(getattr(got_m, key), getattr(want_m, key)), key
