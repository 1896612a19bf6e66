# SPDX-License-Identifier: Apache-2.0
"""ROCm device arrays in and out: stream handles, the fence streams of an input, the intake of
``__cuda_array_interface__`` producers and the ``DeviceArray`` result type (the reference's
src/eda/_device.py:45-334 for the interface: the same protocols, argument meaning and errors, on
HIP devices).

A device input is read in place. PyTorch-ROCm tensors and DLPack producers on a ROCm device go
through ``torch.from_dlpack`` (which runs the producer's own stream handshake); an object that only
exposes ``__cuda_array_interface__`` (version 2 or 3) is wrapped without a copy, and the call waits
for the streams the interface names. A result of a PyTorch tensor is a PyTorch tensor; a result of
any other device input is a ``DeviceArray``, which exports DLPack (device type kDLROCM) and the
array interface (version 3) over the library's memory."""

from __future__ import annotations

import operator
from typing import Any

import numpy as np

KDL_ROCM = 10          # DLPack device type of ROCm memory
NULL_STREAM = 0        # the HIP null stream (the legacy default stream of a CUDA program)
_STREAM_HELP = "an integer HIP stream handle or an object with __cuda_stream__"


def stream_handle(stream: Any) -> int:
    """The integer handle of a ``stream=`` argument: an integer, an object with the stream protocol
    (``__cuda_stream__() -> (version, handle)``), a PyTorch stream (``cuda_stream``) or a CuPy
    stream (``ptr``). 0 (and the reference's legacy handle 1, which HIP has no stream for) is the
    null stream."""
    bad = TypeError(f"eda: stream must be {_STREAM_HELP}, got {type(stream).__name__}")
    if isinstance(stream, (bool, np.bool_)):
        raise bad
    if hasattr(stream, "__index__"):
        value = stream
    elif hasattr(stream, "__cuda_stream__"):
        pair = stream.__cuda_stream__()
        if not isinstance(pair, tuple) or len(pair) != 2:
            raise bad
        value = pair[1]
    elif hasattr(stream, "cuda_stream"):
        value = stream.cuda_stream
    elif hasattr(stream, "ptr"):
        value = stream.ptr
    else:
        raise bad
    if isinstance(value, (bool, np.bool_)):
        raise bad
    try:
        handle = operator.index(value)
    except TypeError:
        raise bad from None
    if handle < 0:
        raise ValueError(f"eda: stream handle {handle} is negative")
    return NULL_STREAM if handle in (0, 1) else handle


def interface_fences(cai: dict, explicit: int | None) -> tuple[int, ...]:
    """The streams an array-interface input waits for: the stream the producer reports (version 3;
    ``None`` means no wait, 0 is refused as version 3 forbids it), or the null stream when the
    interface is older and ``stream=`` is not given; plus ``stream=`` when given."""
    version = int(cai.get("version", 0))
    fences: list[int] = []
    if version >= 3:
        reported = cai.get("stream")
        if reported is not None:
            if isinstance(reported, (bool, np.bool_)) or not hasattr(reported, "__index__"):
                raise ValueError(f"eda: __cuda_array_interface__ stream {reported!r} is not an integer")
            handle = operator.index(reported)
            if handle <= 0:
                raise ValueError(
                    f"eda: __cuda_array_interface__ stream {handle} is not a valid stream (version 3 forbids 0)"
                )
            fences.append(NULL_STREAM if handle == 1 else handle)
    elif explicit is None:
        fences.append(NULL_STREAM)
    if explicit is not None:
        fences.append(explicit)
    return tuple(dict.fromkeys(fences))


def interface_layout(cai: dict) -> tuple[int, tuple[int, ...], tuple[int, ...], np.dtype]:
    """(pointer, shape, strides in elements, dtype) of an array interface; C order when
    ``strides`` is None. A read-only array is accepted (eda only reads it); a mask, or strides that
    are not whole elements, are refused."""
    try:
        ptr, _readonly = cai["data"]
        shape = tuple(int(n) for n in cai["shape"])
        dtype = np.dtype(cai["typestr"])
    except (KeyError, TypeError, ValueError) as exc:
        raise ValueError(f"eda: malformed __cuda_array_interface__: {exc}") from None
    if cai.get("mask") is not None:
        raise ValueError("eda: a masked __cuda_array_interface__ array is not supported")
    item = dtype.itemsize
    raw = cai.get("strides")
    if raw is None:
        strides, acc = [], 1
        for n in reversed(shape):
            strides.append(acc)
            acc *= n
        elems = tuple(reversed(strides))
    else:
        if len(raw) != len(shape):
            raise ValueError("eda: __cuda_array_interface__ strides and shape differ in length")
        if any(int(s) % item for s in raw):
            raise ValueError("eda: __cuda_array_interface__ strides must be whole elements")
        elems = tuple(int(s) // item for s in raw)
    return int(ptr or 0), shape, elems, dtype


def join_fences(fences: tuple[int, ...], device: int, torch: Any) -> tuple[int, Any]:
    """One stream handle that the library's stream may wait for in place of several, and the
    object to keep alive until the call returns: with no fence, the device's current PyTorch stream
    (a wait that orders nothing the producer asked for); with one, that stream; with more, a stream
    of the device that waits (by events, with no device-wide synchronization) for every fence."""
    if not fences:
        return int(torch.cuda.current_stream(device).cuda_stream), None
    if len(fences) == 1:
        return fences[0], None
    joined = torch.cuda.Stream(device=device)
    for h in fences:
        src = torch.cuda.default_stream(device) if h == NULL_STREAM else torch.cuda.ExternalStream(h, device=device)
        ev = torch.cuda.Event()
        ev.record(src)
        joined.wait_event(ev)
    return int(joined.cuda_stream), joined


class DeviceArray:
    """An eda result in ROCm device memory (the result of device input that is not a PyTorch tensor).

    It owns its memory and exposes it without a copy through DLPack (``__dlpack__``,
    ``__dlpack_device__``: device type kDLROCM) and the CUDA array interface (version 3, which ROCm
    consumers read with HIP pointers): ``torch.from_dlpack(y)``, ``cupy.asarray(y)`` (CuPy-ROCm).
    The data is complete when the eda call returns, so any stream may read it. ``shape`` and
    ``strides`` (in bytes) describe the layout like NumPy's; ``device`` is the HIP device ordinal.
    Users do not construct it."""

    __slots__ = ("_owner", "_ptr", "device", "dtype", "shape", "strides")

    def __init__(self, owner: Any) -> None:
        # owner: the PyTorch tensor (a view of the library's output) that holds the memory
        self._owner = owner
        self._ptr = int(owner.data_ptr())
        self.dtype = np.dtype(str(owner.dtype).removeprefix("torch."))
        self.shape = tuple(int(n) for n in owner.shape)
        self.strides = tuple(int(s) * self.dtype.itemsize for s in owner.stride())
        self.device = int(owner.device.index if owner.device.index is not None else 0)

    @property
    def ndim(self) -> int:
        return len(self.shape)

    @property
    def size(self) -> int:
        n = 1
        for extent in self.shape:
            n *= extent
        return n

    @property
    def nbytes(self) -> int:
        return self.size * self.dtype.itemsize

    def _contiguous(self) -> bool:
        acc = self.dtype.itemsize
        for n, s in zip(reversed(self.shape), reversed(self.strides)):
            if n != 1 and s != acc:
                return False
            acc *= n
        return True

    @property
    def __cuda_array_interface__(self) -> dict[str, Any]:
        return {
            "shape": self.shape,
            "typestr": self.dtype.str,
            "data": (self._ptr, False),
            "strides": None if self.size == 0 or self._contiguous() else self.strides,
            "version": 3,
            "stream": None,
        }

    def __dlpack_device__(self) -> tuple[int, int]:
        return (KDL_ROCM, self.device)

    def __dlpack__(self, *, stream: Any = None, max_version: tuple[int, int] | None = None,
                   dl_device: tuple[int, int] | None = None, copy: bool | None = None) -> Any:
        """A DLPack capsule over the memory (versioned when ``max_version >= (1, 0)``). The data is
        complete already, so ``stream`` is ignored; a copy or another device is refused."""
        if copy:
            raise BufferError("eda: a DeviceArray cannot be copied by __dlpack__; copy it with the consumer")
        if dl_device is not None and tuple(dl_device) != (KDL_ROCM, self.device):
            raise BufferError(f"eda: a DeviceArray lives on ({KDL_ROCM}, {self.device}); got dl_device={dl_device}")
        if max_version is None:
            return self._owner.__dlpack__()
        return self._owner.__dlpack__(max_version=max_version)

    def __array__(self, *args: Any, **kwargs: Any) -> np.ndarray:
        raise TypeError(
            "eda: a DeviceArray is ROCm device memory; convert it with its consumer library, for example "
            "torch.from_dlpack(y).cpu()"
        )

    def __repr__(self) -> str:
        return f"DeviceArray(shape={self.shape}, dtype={self.dtype.name}, device={self.device})"
