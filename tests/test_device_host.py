# SPDX-License-Identifier: Apache-2.0
"""The host-side rules of device input (no GPU): stream handles, the fence streams an array
interface names, and its layout (the reference's src/eda/_device.py:45-226 for the protocol
behaviour, on HIP handles)."""

from __future__ import annotations

import numpy as np
import pytest

from eda import _device


class CudaStreamProto:
    def __init__(self, h):
        self.h = h

    def __cuda_stream__(self):
        return (0, self.h)


class TorchLike:
    def __init__(self, h):
        self.cuda_stream = h


class CupyLike:
    def __init__(self, h):
        self.ptr = h


@pytest.mark.parametrize(
    "stream, handle",
    [(5, 5), (np.int64(7), 7), (0, 0), (1, 0), (CudaStreamProto(9), 9), (TorchLike(11), 11), (CupyLike(13), 13)],
)
def test_stream_handles(stream, handle):
    assert _device.stream_handle(stream) == handle


@pytest.mark.parametrize("bad", [True, np.bool_(False), "3", 2.0, object(), CudaStreamProto(True)])
def test_stream_handles_refused(bad):
    with pytest.raises(TypeError, match="eda: stream must be"):
        _device.stream_handle(bad)


def test_negative_stream_handle():
    with pytest.raises(ValueError, match="negative"):
        _device.stream_handle(-1)


def cai(**kw):
    d = {"shape": (4, 5), "typestr": "<f4", "data": (4096, False), "version": 3, "stream": None}
    d.update(kw)
    return d


def test_interface_fences():
    assert _device.interface_fences(cai(), None) == ()
    assert _device.interface_fences(cai(), 7) == (7,)
    assert _device.interface_fences(cai(stream=5), None) == (5,)
    assert _device.interface_fences(cai(stream=5), 7) == (5, 7)
    assert _device.interface_fences(cai(stream=5), 5) == (5,)
    assert _device.interface_fences(cai(stream=1), None) == (0,)  # the legacy handle: the null stream
    assert _device.interface_fences(cai(version=2), None) == (0,)
    assert _device.interface_fences(cai(version=2), 7) == (7,)
    for bad in (0, -3):
        with pytest.raises(ValueError, match="version 3 forbids 0"):
            _device.interface_fences(cai(stream=bad), None)
    with pytest.raises(ValueError, match="not an integer"):
        _device.interface_fences(cai(stream=True), None)


def test_interface_layout():
    ptr, shape, strides, dt = _device.interface_layout(cai())
    assert (ptr, shape, strides, dt) == (4096, (4, 5), (5, 1), np.dtype(np.float32))
    _, _, strides, _ = _device.interface_layout(cai(strides=(4, 16)))
    assert strides == (1, 4)
    _, _, strides, dt = _device.interface_layout(cai(typestr="<f8", strides=(8, 64)))
    assert strides == (1, 8) and dt == np.float64
    with pytest.raises(ValueError, match="whole elements"):
        _device.interface_layout(cai(strides=(6, 20)))
    with pytest.raises(ValueError, match="masked"):
        _device.interface_layout(cai(mask=object()))
    with pytest.raises(ValueError, match="malformed"):
        _device.interface_layout({"shape": (2,)})
