# SPDX-License-Identifier: Apache-2.0
"""``eda.decompose``: argument rules, layout, errors and the result types, over the C ABI of
libeda_hip.so. The rules, the result layout and the error policy follow the reference's
src/eda/_api.py:184-353 (``decompose``), :433-526 (argument checks), :539-607 (layout, info,
errors), so that the call is a drop-in.

Host input (NumPy, or anything ``numpy.asarray`` takes) returns NumPy arrays. ROCm device input is
read in place: a PyTorch-ROCm tensor returns a PyTorch tensor on its device; any other device array
(a DLPack producer on a ROCm device, or an object with ``__cuda_array_interface__``) returns an
``eda.DeviceArray`` (``_device.py``). Results are complete on return."""

from __future__ import annotations

import ctypes
import math
import operator
from collections.abc import Sequence
from dataclasses import dataclass
from typing import Any

import numpy as np

from . import _device, _native
from ._errors import DecompositionError

ERROR_NON_FINITE = 1
ERROR_NON_FINITE_RESULT = 2
ERROR_INTERNAL = 255
_ERROR_MESSAGES = {
    ERROR_NON_FINITE: "the input has inf or NaN; check_finite=True",
    ERROR_NON_FINITE_RESULT: (
        "a non-finite intermediate or output row: the decomposition overflows the dtype's range, "
        "or the input has inf or NaN with check_finite=False"
    ),
}
_FLOAT_DTYPES = ("float32", "float64")
_ERRORS = ("raise", "status")
_INT32_MAX = 2**31 - 1
_KDL_ROCM = 10


@dataclass(frozen=True, eq=False)
class DecompositionInfo:
    """Per-call information of ``eda.decompose(..., return_info=True)`` (the reference's
    src/eda/_api.py:35-61): ``n_imfs``, ``error`` have the batch shape; ``windows``, ``spacing``,
    ``sifts``, ``energy`` the batch shape plus K (the largest IMF count), 0 beyond an element's
    ``n_imfs`` and for a failed element; ``dtype`` the compute dtype; ``axes`` the spatial axes."""

    n_imfs: np.ndarray
    error: np.ndarray
    windows: np.ndarray
    spacing: np.ndarray
    sifts: np.ndarray
    energy: np.ndarray
    dtype: str
    axes: tuple[int, ...]


@dataclass(frozen=True, eq=False)
class Orientation:
    """The orientation of every mode, from ``eda.decompose(..., return_orientation=True)`` (the
    reference's src/eda/_api.py:64-83). With ``m = min(axes)``, ``K`` the largest IMF count of the
    call and ``D`` the number of spatial axes (2 or 3):

    - ``direction``: shape ``x.shape[:m] + (K, D) + x.shape[m:]``, the unit direction ``n`` across
      the stripes that the method computes at the start of each mode; component ``a`` is along the
      ``a``-th axis of ``axes``; its first component of magnitude above the dtype's default tie
      tolerance is positive.
    - ``coherence``: shape ``x.shape[:m] + (K,) + x.shape[m:]``, from 0 (no direction) to 1
      (perfect stripes).

    Rows beyond an element's ``n_imfs``, and all rows of a failed element, are 0. NumPy arrays for
    host input, PyTorch tensors on the input's device for ROCm input, in the compute dtype."""

    direction: Any
    coherence: Any


@dataclass(frozen=True, eq=False)
class Monogenic:
    """The monogenic signal of every field, from ``eda.monogenic`` (the reference's
    src/eda/_api.py:86-108).

    - ``amplitude``, ``phase``, ``frequency``: the shape of ``x``. The phase lies in (-pi, pi]; the
      frequency is in cycles per sample along ``direction``.
    - ``direction``: shape ``x.shape[:m] + (D,) + x.shape[m:]`` with ``m = min(axes)``, the unit
      direction across the local stripes, component ``a`` along the ``a``-th axis of ``axes``, its
      first component of magnitude above the dtype's default tie tolerance positive; ``None`` in
      1-D.
    - ``coherence``: the shape of ``x``, from 0 (no direction) to 1; ``None`` in 1-D.

    NumPy arrays for host input, PyTorch tensors on the input's device for ROCm input, in the
    compute dtype."""

    amplitude: Any
    phase: Any
    frequency: Any
    direction: Any
    coherence: Any


# ----------------------------------------------------------------------------------------------
# Arguments. The accepted types, the clamping and the exception classes and messages are the
# reference's interface (src/eda/_api.py:433-526); the checks themselves are written here.

_MESSAGES = {
    "bool_int": "eda: {name} must be an integer, not bool",
    "not_int": "eda: {name} must be an integer, got {kind}",
    "sifts_range": "eda: sifts must be an integer from 1 to {top}, got {value}",
    "tie_type": "eda: tie_tolerance must be a real number or None, got {kind}",
    "tie_range": "eda: tie_tolerance must be a finite number >= 0, got {value}",
    "flag": "eda: {name} must be True or False, got {kind}",
    "no_axes": "eda: expected an array with at least one dimension",
    "all_axes": (
        "eda: axes=None takes every axis and needs x.ndim <= 3, got {ndim}: name the spatial axes "
        "(1 to 3) with axes="
    ),
    "axes_type": "eda: axes must be an int, a sequence of ints or None, got {kind}",
    "axes_unique": "eda: axes must be unique, got {value}",
    "axes_count": "eda: 1 to 3 spatial axes, got {count}",
}
_BOOLS = (bool, np.bool_)
_REALS = (int, float, np.integer, np.floating)


def _fail(exc: type[Exception], key: str, **fields: Any) -> Exception:
    return exc(_MESSAGES[key].format(**fields))


def _index_of(value: Any, name: str) -> int:
    """``value`` as a Python int through ``__index__`` (bool refused)."""
    if isinstance(value, _BOOLS):
        raise _fail(TypeError, "bool_int", name=name)
    hook = getattr(type(value), "__index__", None)
    if hook is None:
        raise _fail(TypeError, "not_int", name=name, kind=type(value).__name__)
    return int(hook(value))


def _max_imf(value: Any) -> int:
    # clamped into int32: any bound at or below 0 means "no limit", any above 2^31 - 1 is unreachable
    return int(np.clip(_index_of(value, "max_imf"), -(_INT32_MAX + 1), _INT32_MAX))


def _sifts(value: Any) -> int:
    number = _index_of(value, "sifts")
    if number < 1 or number > _INT32_MAX:
        raise _fail(ValueError, "sifts_range", top=_INT32_MAX, value=number)
    return number


def _tie_tolerance(value: Any) -> float | None:
    if value is None:
        return None
    if isinstance(value, _BOOLS) or not isinstance(value, _REALS):
        raise _fail(TypeError, "tie_type", kind=type(value).__name__)
    # an integer too large for a double counts as infinite, which the range check refuses
    too_big = isinstance(value, int) and abs(value) >= 2**1024
    number = math.inf if too_big else float(value)
    if math.isnan(number) or math.isinf(number) or number < 0:
        raise _fail(ValueError, "tie_range", value=value)
    return number


_DCT_MODES = {"exact": 0, "bf16x3": 1}


def _dct_mode(value: Any) -> int:
    """The DCT contraction mode: 'exact' (float32 / float64 transforms) or 'bf16x3' (split-BF16)."""
    if not isinstance(value, str) or value not in _DCT_MODES:
        raise ValueError(f"eda: dct={value!r} is not one of 'exact', 'bf16x3'")
    return _DCT_MODES[value]


def _flag(value: Any, name: str) -> bool:
    if isinstance(value, _BOOLS):
        return bool(value)
    raise _fail(TypeError, "flag", name=name, kind=type(value).__name__)


def _dtype_name(dtype: Any) -> str:
    """The NumPy name of a dtype; a framework dtype (``torch.float32``) by its own name."""
    try:
        return np.dtype(dtype).name
    except TypeError:
        text = str(dtype)
        return text[len("torch."):] if text.startswith("torch.") else text


def _normalize_axes(axes: Any, ndim: int) -> tuple[int, ...]:
    """The spatial axes of an ``ndim``-dimensional input as non-negative ints, in the given order."""
    if ndim < 1:
        raise _fail(ValueError, "no_axes")
    if axes is None:
        if ndim <= 3:
            return tuple(range(ndim))
        raise _fail(ValueError, "all_axes", ndim=ndim)
    one = isinstance(axes, (int, np.integer)) and not isinstance(axes, bool)
    items = [axes] if one else axes
    try:
        given = [operator.index(a) if not isinstance(a, _BOOLS) else None for a in items]
    except TypeError:
        given = [None]
    if None in given:
        raise _fail(TypeError, "axes_type", kind=type(axes).__name__)
    bad = next((a for a in given if a < -ndim or a >= ndim), None)
    if bad is not None:
        raise np.exceptions.AxisError(bad, ndim, "eda")
    wrapped = tuple(a + ndim if a < 0 else a for a in given)
    if len(frozenset(wrapped)) < len(wrapped):
        raise _fail(ValueError, "axes_unique", value=tuple(given))
    if len(wrapped) not in (1, 2, 3):
        raise _fail(ValueError, "axes_count", count=len(wrapped))
    return wrapped


def _restore_order(ndim: int, batch_axes: tuple[int, ...], axes: tuple[int, ...], extra: int = 1) -> list[int]:
    """The permutation (for ``transpose``) from a computed array laid out as ``(*batch, *E,
    *spatial)`` (``extra`` new axes E, the spatial axes in the order of ``axes``) to the public
    layout: the input's axes in their own order with E inserted just before ``min(axes)``."""
    carried = list(batch_axes) + [None] * extra + list(axes)  # the input axis each computed axis holds
    slot = {ax: pos for pos, ax in enumerate(carried) if ax is not None}
    fresh = [pos for pos, ax in enumerate(carried) if ax is None]
    split = min(axes)
    return [slot[f] for f in range(split)] + fresh + [slot[f] for f in range(split, ndim)]


def _error_summary(codes: np.ndarray, indices: np.ndarray, total: int) -> str:
    """One message for the failed elements of a call, grouped by error code (at most 8 indices each)."""
    groups = []
    for code in np.unique(codes.astype(np.int64)):
        hit = indices[codes == code]
        listed = ", ".join(map(str, hit[:8].tolist()))
        if hit.size > 8:
            listed += ", ..."
        reason = _ERROR_MESSAGES.get(int(code), f"error code {int(code)}")
        groups.append(f"code {int(code)} ({reason}) at batch index {listed}")
    head = f"eda: the decomposition failed for {indices.size} of {total} element(s): "
    return head + "; ".join(groups)


# ----------------------------------------------------------------------------------------------
# Input intake


def _torch():
    try:
        import torch  # noqa: PLC0415

        return torch
    except Exception:  # pragma: no cover - torch is optional for host input
        return None


def _is_device_input(x: Any) -> bool:
    torch = _torch()
    if torch is not None and isinstance(x, torch.Tensor):
        return x.device.type == "cuda"
    dev = getattr(x, "__dlpack_device__", None)
    if dev is not None and hasattr(x, "__dlpack__"):
        try:
            return int(dev()[0]) == _KDL_ROCM
        except Exception:
            return False
    return hasattr(x, "__cuda_array_interface__")


def _device_tensor(x: Any):
    torch = _torch()
    if torch is None:
        raise TypeError("eda: device input needs PyTorch-ROCm (the DLPack import of ROCm arrays)")
    if isinstance(x, torch.Tensor):
        return x
    return torch.from_dlpack(x)


def _call_native(desc: _native.ArrayDesc, max_imf: int, tie: float | None, sifts: int, check_finite: bool,
                 stream: int, orientation: bool = False, dct_mode: int = 0) -> _native.Result:
    lib = _native.core()
    p = _native.Params()
    p.max_imf = max_imf
    p.sifts = sifts
    p.check_finite = 1 if check_finite else 0
    p.orientation = 1 if orientation else 0
    p.tie_tolerance = -1.0 if tie is None else tie
    p.max_chunk = 0
    p.stream = stream
    p.dct_mode = dct_mode
    res = _native.Result()
    err = _native.errbuf()
    _native.check(lib.eda_decompose(ctypes.byref(desc), ctypes.byref(p), ctypes.byref(res), err, len(err)), err)
    return res


def _info_of(res: _native.Result, batch_shape: tuple[int, ...], dtype: str, axes: tuple[int, ...]) -> DecompositionInfo:
    B, K = int(res.batch), int(res.K)

    def arr(ptr, n, np_dtype):
        if n == 0:
            return np.zeros(0, np_dtype)
        return np.ctypeslib.as_array(ptr, shape=(n,)).astype(np_dtype, copy=True)

    n_imfs = arr(res.n_imfs, B, np.int32).reshape(batch_shape)
    error = arr(res.error, B, np.int32).reshape(batch_shape)
    shape_k = batch_shape + (K,)
    return DecompositionInfo(
        n_imfs=n_imfs,
        error=error,
        windows=arr(res.windows, B * K, np.int32).reshape(shape_k),
        spacing=arr(res.spacing, B * K, np.float64).reshape(shape_k),
        sifts=arr(res.sifts, B * K, np.int32).reshape(shape_k),
        energy=arr(res.energy, B * K, np.float64).reshape(shape_k),
        dtype=dtype,
        axes=axes,
    )


@dataclass
class _Call:
    """One call's input as the C ABI takes it: the descriptor, the array it points into (kept alive
    until the call returns), the layout and the stream to wait for."""

    desc: _native.ArrayDesc
    keep_alive: Any
    on_device: bool
    target: str
    code: int
    ndim: int
    axes: tuple[int, ...]
    batch_axes: tuple[int, ...]
    spatial: tuple[int, ...]
    batch_shape: tuple[int, ...]
    B: int
    stream_handle: int
    device: Any = None  # the torch.device of device input
    wrap: bool = False  # device input that is not a PyTorch tensor: results are DeviceArrays


def _wrap(call: _Call, t: Any) -> Any:
    return _device.DeviceArray(t) if call.wrap else t


def _device_layout(x: Any, torch: Any) -> tuple:
    """(pointer, shape, strides in elements, dtype name, device ordinal, complex, object to keep
    alive, array-interface dict or None) of a device input: a PyTorch tensor, a DLPack producer on a
    ROCm device (imported by PyTorch), or an object with ``__cuda_array_interface__``."""
    if isinstance(x, torch.Tensor) or (hasattr(x, "__dlpack__") and hasattr(x, "__dlpack_device__")):
        t = _device_tensor(x)
        dev = t.device.index if t.device.index is not None else torch.cuda.current_device()
        return (t.data_ptr() if t.numel() else 0, tuple(t.shape), tuple(t.stride()),
                str(t.dtype).removeprefix("torch."), dev, t.is_complex(), t, None)
    cai = x.__cuda_array_interface__
    ptr, shape, strides, np_dt = _device.interface_layout(cai)
    dev = torch.cuda.current_device()
    if ptr:
        got = ctypes.c_int32(-1)
        _native.core().eda_pointer_device(ctypes.c_void_p(ptr), ctypes.byref(got))
        if got.value < 0:
            raise ValueError("eda: the __cuda_array_interface__ pointer is not ROCm device memory")
        dev = got.value
    return ptr, shape, strides, np_dt.name, dev, np_dt.kind == "c", x, cai


def _prepare(x: Any, axes: Any, dtype: Any, stream: Any) -> _Call:
    """The intake of ``decompose`` and ``monogenic`` (reference src/eda/_api.py:143-181): the dtype
    rules, the axes, the batch layout, host staging or the device array read in place."""
    on_device = _is_device_input(x)
    torch = _torch() if on_device else None
    if on_device:
        if torch is None:
            raise TypeError("eda: device input needs PyTorch-ROCm")
        explicit = None if stream is None else _device.stream_handle(stream)
        dptr, shape, dstrides, src, dev_index, is_complex, owner, cai = _device_layout(x, torch)
        kind_ok = not is_complex
    else:
        data = np.asarray(x)
        src = data.dtype.name
        shape = data.shape
        kind_ok = data.dtype.kind in "biuf"
    if not kind_ok:
        raise TypeError(f"eda: arrays of dtype {src} are not supported (real bool, integer and float arrays only)")
    target = src if dtype is None else _dtype_name(dtype)
    if target not in _FLOAT_DTYPES:
        if dtype is None:
            raise TypeError(
                f"eda: expected a float32 or float64 array, got dtype {src} (pass dtype='float32' "
                "or dtype='float64' to cast it)"
            )
        raise TypeError(f"eda: dtype= must be float32 or float64, got {target}")
    ndim = len(shape)
    axes = _normalize_axes(axes, ndim)
    batch_axes = tuple(i for i in range(ndim) if i not in axes)
    spatial = tuple(shape[a] for a in axes)
    batch_shape = tuple(shape[i] for i in batch_axes)
    B = math.prod(batch_shape)
    order = list(batch_axes) + list(axes)
    code = 0 if target == "float32" else 1

    desc = _native.ArrayDesc()
    desc.dtype = code
    desc.dim = len(axes)
    desc.batch = B
    for a, n in enumerate(spatial):
        desc.n[a] = n
    keep_alive = None
    stream_handle = 0
    if on_device:
        if src not in _native.SOURCE_CODES:
            raise TypeError(f"eda: arrays of dtype {src} are not supported (real bool, integer and float arrays only)")
        # read in place through the strides of every axis, cast to the compute dtype on the fly
        # (reference cpp/src/decompose.cpp:25-82, cpp/src/field/convert.cu)
        pshape = [shape[i] for i in order]
        pstrides = [dstrides[i] for i in order]
        nb = len(batch_axes)
        bshape, bstrides = tuple(pshape[:nb]), tuple(pstrides[:nb])
        device = dev_index
        desc.data = dptr if math.prod(shape) else None
        desc.device = device
        desc.source = 0 if src == target else _native.SOURCE_CODES[src]
        for a in range(len(axes)):
            desc.strides[a] = pstrides[nb + a]
        arrays = None
        if nb == 1:
            desc.batch_stride = bstrides[0]
        elif nb > 1:
            arrays = ((ctypes.c_int64 * nb)(*bshape), (ctypes.c_int64 * nb)(*bstrides))
            desc.batch_axes = nb
            desc.batch_shape = ctypes.cast(arrays[0], ctypes.POINTER(ctypes.c_int64))
            desc.batch_strides = ctypes.cast(arrays[1], ctypes.POINTER(ctypes.c_int64))
        # the streams the input is read after: an array interface names its own (and stream=);
        # tensors and DLPack imports are ordered on the current PyTorch stream (or stream=)
        if cai is not None:
            stream_handle, joined = _device.join_fences(_device.interface_fences(cai, explicit), device, torch)
        else:
            joined = None
            stream_handle = explicit if explicit is not None else int(torch.cuda.current_stream(device).cuda_stream)
        keep_alive = (owner, arrays, joined)
    else:
        xc = data if data.dtype == np.dtype(target) else data.astype(target)
        xt = np.ascontiguousarray(xc.transpose(order)) if order != list(range(ndim)) else np.ascontiguousarray(xc)
        xb = xt.reshape((B,) + spatial)
        keep_alive = xb
        desc.data = xb.ctypes.data if xb.size else None
        desc.device = -1

    return _Call(desc, keep_alive, on_device, target, code, ndim, axes, batch_axes, spatial, batch_shape, B,
                 stream_handle, torch.device("cuda", dev_index) if on_device else None,
                 on_device and not isinstance(x, torch.Tensor))


def decompose(
    x: Any,
    axes: int | Sequence[int] | None = None,
    max_imf: int = -1,
    *,
    dtype: Any = None,
    tie_tolerance: float | None = None,
    sifts: int = 3,
    check_finite: bool = True,
    errors: str = "raise",
    stream: Any = None,
    return_info: bool = False,
    return_orientation: bool = False,
    dct: str = "exact",
) -> Any:
    """Decompose every 1-D, 2-D or 3-D field of ``x`` into intrinsic mode functions (IMFs) and a
    residue, on an AMD Instinct GPU (gfx950).

    The arguments, the result layout ``x.shape[:m] + (P,) + x.shape[m:]`` with ``m = min(axes)``,
    the information and the error policy are those of the reference's ``eda.decompose``: each batch
    element is decomposed on its own; along the mode axis each element holds its IMFs from fine to
    coarse, then its residue, then zero rows; ``P`` is one more than the largest IMF count over the
    elements without an error. Every spatial extent must be at least 2; 1-D signals may hold up to
    2^18 samples, fields of 2 or 3 axes up to 8192 (float32) or 4096 (float64) per axis (the
    reference stops at 8192 / 4096 for every field). ``return_orientation=True`` (2 or 3 spatial
    axes) also returns an ``Orientation``: the direction and the coherence that the method
    computes at the start of each mode. The result is ``y``, then ``info`` if ``return_info``, then the ``Orientation`` if
    ``return_orientation``; asking for the orientation does not change ``y`` or ``info``.
    ``dct="bf16x3"`` (float32 only; an addition to the reference's arguments) runs every DCT and
    Riesz contraction of the orientation as split-BF16 products on the matrix cores (three bf16
    products per term, float32 accumulation) instead of the exact float32 transforms.

    Host input returns NumPy arrays. A PyTorch-ROCm tensor returns a PyTorch tensor on its device;
    any other ROCm device array (a DLPack producer on a ROCm device, or an object with
    ``__cuda_array_interface__``) returns an ``eda.DeviceArray``. The call reads device input after
    the current PyTorch stream of the device, or after ``stream`` (an integer HIP stream handle, an
    object with ``__cuda_stream__``, a ``torch.cuda.Stream``); an array-interface input is read
    after the stream it reports (version 3; older versions: the null stream) and ``stream``. The
    result is complete on return."""
    max_imf = _max_imf(max_imf)
    tie = _tie_tolerance(tie_tolerance)
    sifts = _sifts(sifts)
    check_finite = _flag(check_finite, "check_finite")
    return_info = _flag(return_info, "return_info")
    return_orientation = _flag(return_orientation, "return_orientation")
    if not isinstance(errors, str) or errors not in _ERRORS:
        raise ValueError(f"eda: errors={errors!r} is not one of 'raise', 'status'")
    dct_mode = _dct_mode(dct)

    call = _prepare(x, axes, dtype, stream)
    if return_orientation and len(call.axes) < 2:
        raise ValueError(
            "eda: return_orientation=True needs 2 or 3 spatial axes, got 1 (the method has no "
            "orientation in 1-D)"
        )
    desc, keep_alive, on_device, target, code = call.desc, call.keep_alive, call.on_device, call.target, call.code
    ndim, axes, batch_axes, spatial, batch_shape, B = (call.ndim, call.axes, call.batch_axes, call.spatial,
                                                       call.batch_shape, call.B)
    stream_handle = call.stream_handle
    tdevice = call.device
    wrap = _device.DeviceArray if call.wrap else (lambda t: t)
    out_final = _restore_order(ndim, batch_axes, axes)
    del call
    res = _call_native(desc, max_imf, tie, sifts, check_finite, stream_handle, return_orientation, dct_mode)
    try:
        info = _info_of(res, batch_shape, target, axes)
        codes = info.error.reshape(-1)
        internal = np.flatnonzero(codes == ERROR_INTERNAL)
        if internal.size:
            raise RuntimeError(
                "eda: an internal invariant failed (error code 255) for batch index "
                f"{', '.join(str(int(i)) for i in internal[:8])}: this is a bug in eda"
            )
        failed = np.flatnonzero(codes != 0)
        if failed.size and errors == "raise":
            raise DecompositionError(
                _error_summary(codes[failed], failed, codes.size), indices=failed, codes=codes[failed], info=info
            )
        P = int(res.rows)
        lib = _native.core()
        err = _native.errbuf()
        if on_device:
            torch = _torch()
            out = torch.empty((B, P) + spatial, dtype=torch.float32 if code == 0 else torch.float64,
                              device=tdevice)
            if out.numel():
                _native.check(lib.eda_collect(ctypes.byref(res), ctypes.c_void_p(out.data_ptr()), int(desc.device),
                                              err, len(err)), err)
            y = wrap(out.reshape(batch_shape + (P,) + spatial).permute(out_final))
        else:
            out = np.empty((B, P) + spatial, dtype=np.dtype(target))
            if out.size:
                _native.check(lib.eda_collect(ctypes.byref(res), out.ctypes.data_as(ctypes.c_void_p), -1, err,
                                              len(err)), err)
            computed = out.reshape(batch_shape + (P,) + spatial)
            y = computed.transpose(out_final) if out_final != list(range(ndim + 1)) else computed
        orient = None
        if return_orientation:
            orient = _orientation(res, batch_shape, spatial, ndim, batch_axes, axes, target,
                                  tdevice if on_device else None, wrap)
    finally:
        _native.core().eda_release(ctypes.byref(res))
    del keep_alive
    out_t: tuple = (y,)
    if return_info:
        out_t += (info,)
    if return_orientation:
        out_t += (orient,)
    return out_t if len(out_t) > 1 else y


def _orientation(res: _native.Result, batch_shape: tuple[int, ...], spatial: tuple[int, ...], ndim: int,
                 batch_axes: tuple[int, ...], axes: tuple[int, ...], target: str, device: Any,
                 wrap: Any = None) -> Orientation:
    """The ``Orientation`` of a call from the kept direction (B, K, D, *spatial) and coherence (B,
    K, *spatial), in the public layout (reference src/eda/_api.py:558-577)."""
    D, K = len(axes), int(res.K)
    B = math.prod(batch_shape)
    shapes = (batch_shape + (K, D) + spatial, batch_shape + (K,) + spatial)
    perms = [_restore_order(ndim, batch_axes, axes, e) for e in (2, 1)]
    lib = _native.core()
    err = _native.errbuf()
    if device is not None:
        torch = _torch()
        tdt = torch.float32 if target == "float32" else torch.float64
        dirs = torch.empty((B, K, D) + spatial, dtype=tdt, device=device)
        coh = torch.empty((B, K) + spatial, dtype=tdt, device=device)
        if coh.numel():
            _native.check(lib.eda_collect_orientation(ctypes.byref(res), ctypes.c_void_p(dirs.data_ptr()),
                                                      ctypes.c_void_p(coh.data_ptr()), int(device.index), err,
                                                      len(err)), err)
        arrays = [a.reshape(sh).permute(pm) for a, sh, pm in zip((dirs, coh), shapes, perms)]
        return Orientation(*(wrap(a) for a in arrays) if wrap is not None else arrays)
    dirs = np.empty((B, K, D) + spatial, dtype=np.dtype(target))
    coh = np.empty((B, K) + spatial, dtype=np.dtype(target))
    if coh.size:
        _native.check(lib.eda_collect_orientation(ctypes.byref(res), dirs.ctypes.data_as(ctypes.c_void_p),
                                                  coh.ctypes.data_as(ctypes.c_void_p), -1, err, len(err)), err)
    arrays = []
    for a, sh, pm in zip((dirs, coh), shapes, perms):
        a = a.reshape(sh)
        arrays.append(a.transpose(pm) if pm != list(range(len(pm))) else a)
    return Orientation(*arrays)


def monogenic(
    x: Any,
    axes: int | Sequence[int] | None = None,
    *,
    dtype: Any = None,
    check_finite: bool = True,
    stream: Any = None,
) -> Monogenic:
    """The monogenic signal of every 1-D, 2-D or 3-D field of ``x`` on an AMD Instinct GPU (gfx950):
    its local amplitude, phase, frequency and, in 2-D and 3-D, the direction and the coherence of
    the local stripes (reference src/eda/_api.py:356-426).

    ``x``, ``axes``, ``dtype``, ``check_finite`` and ``stream`` follow the rules of
    ``eda.decompose``. Each field is processed on its own, and its result is the same, bit for bit,
    alone or at any position of any batch. For a field f with the half-sample symmetric extension
    at the borders, the first-order Riesz transform R1 = (R_a f), the second-order one R2 = (R_a R_b
    f) and the direction n (the top eigenvector of T = R1 R1^T + R2 R2^T in 2-D and 3-D, 1 in 1-D),
    with q = n . R1: amplitude = sqrt(f^2 + |R1|^2); phase = atan2(q, f) in (-pi, pi]; frequency =
    (f dq/dn - q df/dn) / (2 pi amplitude^2) in cycles per sample along n; coherence = (mu1 - mu2)
    / (mu1 + mu2) from the two largest eigenvalues of T. The phase and the frequency are 0 where the
    amplitude vanishes. With ``check_finite=True`` a field that holds inf or NaN raises
    ``ValueError``."""
    check_finite = _flag(check_finite, "check_finite")
    call = _prepare(x, axes, dtype, stream)
    D, B = len(call.axes), call.B
    count = math.prod(call.spatial)
    planes = 3 if D == 1 else 4 + D
    n = B * count
    lib = _native.core()
    err = _native.errbuf()
    nonfinite = np.zeros(max(B, 1), np.int32)
    chunk = ctypes.c_int64(0)
    if call.on_device:
        torch = _torch()
        tdt = torch.float32 if call.code == 0 else torch.float64
        out = torch.empty(planes * n, dtype=tdt, device=call.device)
        ptr, out_device = (out.data_ptr() if n else None), int(call.desc.device)
    else:
        out = np.empty(planes * n, dtype=np.dtype(call.target))
        ptr, out_device = (out.ctypes.data if n else None), -1
    _native.check(lib.eda_monogenic(ctypes.byref(call.desc), 1 if check_finite else 0, 0, call.stream_handle,
                                    ctypes.c_void_p(ptr), out_device,
                                    nonfinite.ctypes.data_as(ctypes.POINTER(ctypes.c_int32)), ctypes.byref(chunk),
                                    err, len(err)), err)
    bad = np.flatnonzero(nonfinite[:B])
    if check_finite and bad.size:
        shown = ", ".join(str(int(i)) for i in bad[:8]) + (", ..." if bad.size > 8 else "")
        raise ValueError(
            f"eda: the input has inf or NaN at batch index {shown} ({bad.size} of {B} field(s)); check_finite=True"
        )
    # the planes: amplitude, phase, frequency, then in 2-D and 3-D the direction (D components per
    # sample) and the coherence
    parts, offset = [], 0
    for comps in ((0, 0, 0, D, 0) if D >= 2 else (0, 0, 0)):
        extra = (comps,) if comps else ()
        computed = call.batch_shape + extra + call.spatial
        perm = _restore_order(call.ndim, call.batch_axes, call.axes, len(extra))
        size = n * max(comps, 1)
        part = out[offset:offset + size]
        if call.on_device:
            parts.append(_wrap(call, part.view(computed).permute(perm)))
        else:
            part = part.reshape(computed)
            parts.append(part.transpose(perm) if perm != list(range(len(perm))) else part)
        offset += size
    del call
    return Monogenic(*parts) if D >= 2 else Monogenic(*parts, None, None)
