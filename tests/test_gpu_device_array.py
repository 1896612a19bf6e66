# SPDX-License-Identifier: Apache-2.0
"""Device arrays in and out (GPU): a result of device input that is not a PyTorch tensor is an
eda.DeviceArray that exports DLPack (kDLROCM) and the array interface over the library's memory;
inputs that only expose DLPack or only ``__cuda_array_interface__`` (version 2 and 3, contiguous and
strided) are read in place, bitwise as the PyTorch tensor is; an array-interface input is read
after the stream it reports and after ``stream=`` (event waits, several fences at once)."""

from __future__ import annotations

import dataclasses

import numpy as np
import pytest

import eda
import workloads

pytestmark = pytest.mark.gpu


def torch_cuda():
    import torch

    assert torch.cuda.is_available()
    return torch


class DLPackOnly:
    """A producer that exposes only DLPack (what a non-PyTorch ROCm library hands over)."""

    def __init__(self, t):
        self.t = t

    def __dlpack__(self, **kw):
        return self.t.__dlpack__(**kw)

    def __dlpack_device__(self):
        return self.t.__dlpack_device__()


class InterfaceOnly:
    """A producer that exposes only __cuda_array_interface__ over a tensor's memory."""

    def __init__(self, t, version=3, stream=None, strided=True):
        self.t = t
        item = t.element_size()
        self.cai = {
            "shape": tuple(t.shape),
            "typestr": np.dtype(str(t.dtype).removeprefix("torch.")).str,
            "data": (t.data_ptr(), False),
            "strides": tuple(s * item for s in t.stride()) if strided else None,
            "version": version,
        }
        if version >= 3:
            self.cai["stream"] = stream

    @property
    def __cuda_array_interface__(self):
        return self.cai


def images(torch, B=3, n=96, dtype=np.float32):
    return torch.from_numpy(workloads.rings_stripes_blobs(B, n, dtype)).to("cuda")


def bits(t):
    return np.ascontiguousarray(t.detach().cpu().numpy()).view(np.uint8)


def test_dlpack_producer_returns_a_device_array():
    torch = torch_cuda()
    x = images(torch)
    ref, info_ref = eda.decompose(x, axes=(1, 2), return_info=True)
    y, info = eda.decompose(DLPackOnly(x), axes=(1, 2), return_info=True)
    assert isinstance(y, eda.DeviceArray) and not isinstance(ref, eda.DeviceArray)
    assert y.shape == tuple(ref.shape) and y.dtype == np.float32 and y.device == x.device.index
    assert y.ndim == ref.ndim and y.size == ref.numel() and y.nbytes == ref.numel() * 4
    assert y.__dlpack_device__() == (10, x.device.index)
    back = torch.from_dlpack(y)
    assert np.array_equal(bits(back), bits(ref))
    assert np.array_equal(info.n_imfs, info_ref.n_imfs)
    cai = y.__cuda_array_interface__
    assert cai["version"] == 3 and cai["stream"] is None and cai["data"][0] == back.data_ptr()
    assert cai["strides"] is None or len(cai["strides"]) == y.ndim
    with pytest.raises(TypeError, match="DeviceArray is ROCm device memory"):
        np.asarray(y)
    with pytest.raises(BufferError):
        y.__dlpack__(copy=True)
    with pytest.raises(BufferError):
        y.__dlpack__(dl_device=(1, 0))
    assert "DeviceArray(shape=" in repr(y)


@pytest.mark.parametrize("version", [2, 3])
@pytest.mark.parametrize("strided", [False, True])
def test_interface_producer_is_read_in_place(version, strided):
    torch = torch_cuda()
    x = images(torch)
    ref = eda.decompose(x, axes=(1, 2))
    y = eda.decompose(InterfaceOnly(x, version=version, strided=strided), axes=(1, 2))
    assert isinstance(y, eda.DeviceArray)
    assert np.array_equal(bits(torch.from_dlpack(y)), bits(ref))
    # a transposed view through the interface's strides (batch axis last)
    xt = x.permute(1, 2, 0)
    ref_t = eda.decompose(xt, axes=(0, 1))
    y_t = eda.decompose(InterfaceOnly(xt, version=version), axes=(0, 1))
    assert np.array_equal(bits(torch.from_dlpack(y_t)), bits(ref_t))


def test_interface_stream_and_explicit_stream_are_both_waited_for():
    """The input is written on one side stream (the interface's) after a long kernel there, and
    stream= names a second side stream; the call must read the finished input."""
    torch = torch_cuda()
    x = images(torch, B=4, n=128)
    ref = eda.decompose(x, axes=(1, 2))
    s1, s2 = torch.cuda.Stream(), torch.cuda.Stream()
    buf = torch.zeros_like(x)
    big = torch.randn(4096, 4096, device="cuda")
    torch.cuda.synchronize()
    with torch.cuda.stream(s1):
        for _ in range(4):
            big = big @ big * 1e-3  # keeps s1 busy before the copy lands
        buf.copy_(x)
    with torch.cuda.stream(s2):
        torch.cuda._sleep(1000)
    y = eda.decompose(InterfaceOnly(buf, stream=int(s1.cuda_stream)), axes=(1, 2), stream=s2)
    assert np.array_equal(bits(torch.from_dlpack(y)), bits(ref))


def test_monogenic_and_orientation_of_a_producer_are_device_arrays():
    torch = torch_cuda()
    x = images(torch)
    m_ref = eda.monogenic(x, axes=(1, 2))
    m = eda.monogenic(DLPackOnly(x), axes=(1, 2))
    _, o_ref = eda.decompose(x, axes=(1, 2), return_orientation=True)
    _, o = eda.decompose(DLPackOnly(x), axes=(1, 2), return_orientation=True)
    for got, want in ((m, m_ref), (o, o_ref)):
        for f in dataclasses.fields(want):
            a, b = getattr(got, f.name), getattr(want, f.name)
            assert isinstance(a, eda.DeviceArray), f.name
            assert np.array_equal(bits(torch.from_dlpack(a)), bits(b)), f.name
