# SPDX-License-Identifier: Apache-2.0
"""The reference's native-module interface, served by libeda_hip.so.

The reference's Python package calls its nanobind module ``eda._core`` (signatures in
src/eda/_core.pyi:9-27; implementation cpp/python/decompose.cpp:325-509, module.cpp:64-142). This
module exposes the functions of that interface that the decomposition path uses, with the same
names, argument order, argument meaning and results, over the C ABI of include/eda_hip.h, so that
the reference's ``_native.core()`` can return it in place of ``eda._core`` on a ROCm machine:

  decompose_host(x, max_imf, tie_tolerance, sifts, check_finite, orientation, max_chunk)
      -> (rows (B, P, *spatial) NumPy array, info dict)
  decompose_device(ptr, device, dtype, target, batch_shape, batch_strides, shape, strides, fences,
                   max_imf, tie_tolerance, sifts, check_finite, orientation, max_chunk)
      -> (owner, out_ptr, info dict); rows C-contiguous (B, P, *spatial) on ``device``
  monogenic_host(x, check_finite, max_chunk) -> (out flat planes NumPy array, info dict)
  monogenic_device(ptr, device, dtype, target, batch_shape, batch_strides, shape, strides, fences,
                   check_finite, max_chunk) -> (owner, out_ptr, info dict)
  build_info(), device_count(), device_info(device=0), pointer_device(ptr)

info holds ``n_imfs``, ``error`` [B] int32; ``windows``, ``sifts`` [B, K] int32; ``spacing``,
``energy`` [B, K] float64; ``chunk`` int; with ``orientation``, for host input ``direction`` (B, K,
D, *spatial) and ``coherence`` (B, K, *spatial) arrays, for device input ``orientation`` = (owner,
ptr) of one buffer holding the direction followed by the coherence ((None, 0) when B or K is 0).
Errors raise the reference's classes: ValueError, MemoryError, ``eda.CudaError``, RuntimeError.
The monogenic calls' info holds ``nonfinite`` [B] int32 and ``chunk``; their output holds the planes
amplitude, phase, frequency (B, *spatial) each, then for 2 or 3 spatial axes the direction (B, D,
*spatial) and the coherence (B, *spatial). Device input of any of the 13 element types
(bool, int8..int64, uint8..uint64, float16, bfloat16, float32, float64) and any batch strides is
read in place and cast to ``target`` on the device. Not served: the DLPack helpers.
The device path allocates its output through PyTorch-ROCm, whose tensor is the ``owner``."""

from __future__ import annotations

import ctypes
import json
import math
from collections.abc import Sequence

import numpy as np

from . import _native

_CODES = {"float32": 0, "float64": 1}


def build_info() -> dict:
    buf = ctypes.create_string_buffer(4096)
    _native.core().eda_build_info(buf, len(buf))
    return json.loads(buf.value.decode())


def device_count() -> int:
    return int(_native.core().eda_device_count())


def device_info(device: int = 0) -> dict:
    buf = ctypes.create_string_buffer(4096)
    err = _native.errbuf()
    rc = _native.core().eda_device_info(int(device), buf, len(buf))
    if rc != _native.EDA_OK:
        err.value = buf.value
        _native.check(rc, err)
    return json.loads(buf.value.decode())


def pointer_device(ptr: int) -> int:
    dev = ctypes.c_int32(-1)
    rc = _native.core().eda_pointer_device(ctypes.c_void_p(int(ptr)), ctypes.byref(dev))
    if rc != _native.EDA_OK:
        raise ValueError(f"eda: {ptr:#x} is not a pointer the HIP runtime knows")
    return int(dev.value)


def _params(max_imf, tie_tolerance, sifts, check_finite, orientation, max_chunk, stream=0) -> _native.Params:
    if int(sifts) < 1:
        raise ValueError(f"eda: sifts must be at least 1, got {sifts}")
    if int(max_chunk) < 0:
        raise ValueError(f"eda: max_chunk must be >= 0, got {max_chunk}")
    p = _native.Params()
    p.max_imf = int(max_imf)
    p.sifts = int(sifts)
    p.check_finite = 1 if check_finite else 0
    p.orientation = 1 if orientation else 0
    p.tie_tolerance = -1.0 if tie_tolerance is None else float(tie_tolerance)
    p.max_chunk = int(max_chunk)
    p.stream = int(stream)
    return p


def _info(res: _native.Result) -> dict:
    B, K = int(res.batch), int(res.K)

    def arr(ptr, n, dt, shape):
        if n == 0:
            return np.zeros(shape, dt)
        return np.ctypeslib.as_array(ptr, shape=(n,)).astype(dt, copy=True).reshape(shape)

    return {
        "n_imfs": arr(res.n_imfs, B, np.int32, (B,)),
        "error": arr(res.error, B, np.int32, (B,)),
        "windows": arr(res.windows, B * K, np.int32, (B, K)),
        "sifts": arr(res.sifts, B * K, np.int32, (B, K)),
        "spacing": arr(res.spacing, B * K, np.float64, (B, K)),
        "energy": arr(res.energy, B * K, np.float64, (B, K)),
        "chunk": int(res.chunk),
    }


def _run(desc: _native.ArrayDesc, p: _native.Params) -> _native.Result:
    res = _native.Result()
    err = _native.errbuf()
    _native.check(_native.core().eda_decompose(ctypes.byref(desc), ctypes.byref(p), ctypes.byref(res), err,
                                               len(err)), err)
    return res


def decompose_host(x, max_imf: int, tie_tolerance: float | None, sifts: int, check_finite: bool,
                   orientation: bool, max_chunk: int) -> tuple:
    """The C-contiguous host array x (B, *spatial), float32 or float64: (rows (B, P, *spatial),
    info) (reference cpp/python/decompose.cpp:325-384)."""
    x = np.asarray(x)
    if x.dtype.name not in _CODES:
        raise TypeError(f"eda: decompose_host needs float32 or float64, got {x.dtype}")
    if not x.flags.c_contiguous:
        raise ValueError("eda: decompose_host needs a C-contiguous array")
    if not 2 <= x.ndim <= 4:
        raise ValueError(f"eda: expected (B, *spatial) with 1 to 3 spatial axes, got {x.ndim} axes")
    p = _params(max_imf, tie_tolerance, sifts, check_finite, orientation, max_chunk)
    desc = _native.ArrayDesc()
    desc.dtype = _CODES[x.dtype.name]
    desc.dim = x.ndim - 1
    desc.batch = x.shape[0]
    for a, n in enumerate(x.shape[1:]):
        desc.n[a] = n
    desc.device = -1
    desc.data = x.ctypes.data if x.size else None
    res = _run(desc, p)
    try:
        info = _info(res)
        rows = np.empty((x.shape[0], int(res.rows)) + x.shape[1:], dtype=x.dtype)
        if rows.size:
            err = _native.errbuf()
            _native.check(_native.core().eda_collect(ctypes.byref(res), rows.ctypes.data_as(ctypes.c_void_p), -1,
                                                     err, len(err)), err)
        if orientation:
            K = int(res.K)
            dirs = np.zeros((x.shape[0], K, desc.dim) + x.shape[1:], dtype=x.dtype)
            coh = np.zeros((x.shape[0], K) + x.shape[1:], dtype=x.dtype)
            if coh.size:
                err = _native.errbuf()
                _native.check(_native.core().eda_collect_orientation(
                    ctypes.byref(res), dirs.ctypes.data_as(ctypes.c_void_p), coh.ctypes.data_as(ctypes.c_void_p), -1,
                    err, len(err)), err)
            info["direction"] = dirs
            info["coherence"] = coh
    finally:
        _native.core().eda_release(ctypes.byref(res))
    return rows, info


def _merged_batch_axes(batch_shape: Sequence[int], batch_strides: Sequence[int]) -> tuple[list[int], list[int]]:
    """The batch axes without the extent-1 ones, neighbours whose strides chain merged into one
    (reference cpp/src/decompose.cpp:48-59)."""
    shape: list[int] = []
    stride: list[int] = []
    for n, st in zip(batch_shape, batch_strides):
        n, st = int(n), int(st)
        if n == 1:
            continue
        if shape and stride[-1] == st * n:
            shape[-1] *= n
            stride[-1] = st
        else:
            shape.append(n)
            stride.append(st)
    return shape, stride


def _device_desc(ptr: int, device: int, dtype: str, target: str, batch_shape: Sequence[int],
                 batch_strides: Sequence[int], shape: Sequence[int], strides: Sequence[int]):
    """The descriptor of a device array read in place (element strides of its own dtype, any
    sign), cast to ``target`` on the fly; returns it and the ctypes arrays it points to."""
    if target not in _CODES:
        raise ValueError(f"eda: target must be float32 or float64, got {target}")
    if dtype not in _native.SOURCE_CODES:
        raise TypeError(f"eda: arrays of dtype {dtype} are not supported (real bool, integer and float arrays only)")
    if not 1 <= len(shape) <= 3 or len(strides) != len(shape):
        raise ValueError("eda: expected 1 to 3 spatial axes with one stride each")
    if len(batch_shape) != len(batch_strides):
        raise ValueError("eda: batch shape and strides differ in length")
    if len(batch_shape) > 64:
        raise ValueError("eda: a device array may have at most 64 batch axes")
    B = math.prod(int(n) for n in batch_shape)
    count = math.prod(int(n) for n in shape)
    desc = _native.ArrayDesc()
    desc.dtype = _CODES[target]
    desc.source = 0 if dtype == target else _native.SOURCE_CODES[dtype]
    desc.dim = len(shape)
    desc.batch = B
    for a, (n, st) in enumerate(zip(shape, strides)):
        desc.n[a] = int(n)
        desc.strides[a] = int(st)
    mshape, mstride = _merged_batch_axes(batch_shape, batch_strides)
    keep = None
    if len(mshape) == 1:
        desc.batch_stride = mstride[0]
    elif len(mshape) > 1:
        keep = ((ctypes.c_int64 * len(mshape))(*mshape), (ctypes.c_int64 * len(mshape))(*mstride))
        desc.batch_axes = len(mshape)
        desc.batch_shape = ctypes.cast(keep[0], ctypes.POINTER(ctypes.c_int64))
        desc.batch_strides = ctypes.cast(keep[1], ctypes.POINTER(ctypes.c_int64))
    desc.device = int(device)
    desc.data = int(ptr) if B and count else None
    return desc, keep


def _fence_stream(fences: Sequence[int], device: int, torch) -> int:
    """The one stream handle the call waits for by event (0: none); the legacy and per-thread
    default streams (1, 2) and several fences make it wait for the whole device instead."""
    fences = [int(f) for f in fences]
    if not fences:
        return 0
    if len(fences) > 1 or fences[0] in (1, 2):
        torch.cuda.synchronize(int(device))
        return 0
    return fences[0]


def decompose_device(ptr: int, device: int, dtype: str, target: str, batch_shape: Sequence[int],
                     batch_strides: Sequence[int], shape: Sequence[int], strides: Sequence[int],
                     fences: Sequence[int], max_imf: int, tie_tolerance: float | None, sifts: int,
                     check_finite: bool, orientation: bool, max_chunk: int) -> tuple:
    """A ROCm device array (strides in elements, any sign) after waiting for the fence streams:
    (owner, out_ptr, info), the rows C-contiguous (B, P, *spatial) in ``target`` on ``device``,
    complete on return (reference cpp/python/decompose.cpp:452-509)."""
    import torch  # noqa: PLC0415 - the device path allocates its output through PyTorch-ROCm

    desc, keep = _device_desc(ptr, device, dtype, target, batch_shape, batch_strides, shape, strides)
    B = int(desc.batch)
    stream = _fence_stream(fences, device, torch)
    p = _params(max_imf, tie_tolerance, sifts, check_finite, orientation, max_chunk, stream)
    res = _run(desc, p)
    del keep
    try:
        info = _info(res)
        tdt = torch.float32 if target == "float32" else torch.float64
        owner = torch.empty((B, int(res.rows)) + tuple(int(n) for n in shape), dtype=tdt,
                            device=torch.device("cuda", int(device)))
        if owner.numel():
            err = _native.errbuf()
            _native.check(_native.core().eda_collect(ctypes.byref(res), ctypes.c_void_p(owner.data_ptr()),
                                                     int(device), err, len(err)), err)
        if orientation:
            info["orientation"] = (None, 0)
            K, count = int(res.K), math.prod(int(n) for n in shape)
            nc = B * K * count
            if nc:
                buf = torch.empty(nc * (len(shape) + 1), dtype=tdt, device=torch.device("cuda", int(device)))
                err = _native.errbuf()
                item = buf.element_size()
                _native.check(_native.core().eda_collect_orientation(
                    ctypes.byref(res), ctypes.c_void_p(buf.data_ptr()),
                    ctypes.c_void_p(buf.data_ptr() + nc * len(shape) * item), int(device), err, len(err)), err)
                info["orientation"] = (buf, buf.data_ptr())
    finally:
        _native.core().eda_release(ctypes.byref(res))
    return owner, owner.data_ptr(), info


def _monogenic_planes(dim: int) -> int:
    return 3 if dim == 1 else 4 + dim


def monogenic_host(x, check_finite: bool, max_chunk: int) -> tuple:
    """The monogenic signal of the C-contiguous host array x (B, *spatial), float32 or float64: (out,
    info), out the flat planes (reference cpp/python/monogenic.cpp, src/eda/_core.pyi:25-28)."""
    x = np.asarray(x)
    if x.dtype.name not in _CODES:
        raise TypeError(f"eda: monogenic_host needs float32 or float64, got {x.dtype}")
    if not x.flags.c_contiguous:
        raise ValueError("eda: monogenic_host needs a C-contiguous array")
    if not 2 <= x.ndim <= 4:
        raise ValueError(f"eda: expected (B, *spatial) with 1 to 3 spatial axes, got {x.ndim} axes")
    desc = _native.ArrayDesc()
    desc.dtype = _CODES[x.dtype.name]
    desc.dim = x.ndim - 1
    desc.batch = x.shape[0]
    for a, n in enumerate(x.shape[1:]):
        desc.n[a] = n
    desc.device = -1
    desc.data = x.ctypes.data if x.size else None
    out = np.empty(_monogenic_planes(desc.dim) * x.size, dtype=x.dtype)
    nonfinite = np.zeros(max(x.shape[0], 1), np.int32)
    chunk = ctypes.c_int64(0)
    err = _native.errbuf()
    _native.check(_native.core().eda_monogenic(
        ctypes.byref(desc), 1 if check_finite else 0, int(max_chunk), 0,
        ctypes.c_void_p(out.ctypes.data if out.size else None), -1,
        nonfinite.ctypes.data_as(ctypes.POINTER(ctypes.c_int32)), ctypes.byref(chunk), err, len(err)), err)
    return out, {"nonfinite": nonfinite[:x.shape[0]].copy(), "chunk": int(chunk.value)}


def monogenic_device(ptr: int, device: int, dtype: str, target: str, batch_shape: Sequence[int],
                     batch_strides: Sequence[int], shape: Sequence[int], strides: Sequence[int],
                     fences: Sequence[int], check_finite: bool, max_chunk: int) -> tuple:
    """The monogenic signal of a ROCm device array after waiting for the fence streams: (owner,
    out_ptr, info), the planes of monogenic_host in one device buffer, complete on return."""
    import torch  # noqa: PLC0415 - the device path allocates its output through PyTorch-ROCm

    desc, keep = _device_desc(ptr, device, dtype, target, batch_shape, batch_strides, shape, strides)
    B = int(desc.batch)
    count = math.prod(int(n) for n in shape)
    stream = _fence_stream(fences, device, torch)
    tdt = torch.float32 if target == "float32" else torch.float64
    owner = torch.empty(_monogenic_planes(len(shape)) * B * count, dtype=tdt, device=torch.device("cuda", int(device)))
    nonfinite = np.zeros(max(B, 1), np.int32)
    chunk = ctypes.c_int64(0)
    err = _native.errbuf()
    _native.check(_native.core().eda_monogenic(
        ctypes.byref(desc), 1 if check_finite else 0, int(max_chunk), stream,
        ctypes.c_void_p(owner.data_ptr() if owner.numel() else None), int(device),
        nonfinite.ctypes.data_as(ctypes.POINTER(ctypes.c_int32)), ctypes.byref(chunk), err, len(err)), err)
    del keep
    return owner, owner.data_ptr(), {"nonfinite": nonfinite[:B].copy(), "chunk": int(chunk.value)}
