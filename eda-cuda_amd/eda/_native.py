# SPDX-License-Identifier: Apache-2.0
"""The ctypes binding of libeda_hip.so (include/eda_hip.h), loaded on first use.

Replaces the reference's nanobind module eda._core (cpp/python/module.cpp, decompose.cpp) and its
loader (src/eda/_native.py:62-86). There is no CPU fallback: without the built library every call
that needs it raises."""

from __future__ import annotations

import ctypes
import threading
from pathlib import Path

LIB_PATH = Path(__file__).resolve().parent / "libeda_hip.so"

EDA_OK, EDA_EINVAL, EDA_ENOMEM, EDA_EHIP, EDA_EINTERNAL = 0, 1, 2, 3, 4


class ArrayDesc(ctypes.Structure):
    _fields_ = [
        ("data", ctypes.c_void_p),
        ("device", ctypes.c_int32),
        ("dtype", ctypes.c_int32),
        ("dim", ctypes.c_int32),
        ("source", ctypes.c_int32),
        ("batch", ctypes.c_int64),
        ("n", ctypes.c_int64 * 3),
        ("batch_stride", ctypes.c_int64),
        ("strides", ctypes.c_int64 * 3),
        ("batch_axes", ctypes.c_int32),
        ("reserved", ctypes.c_int32),
        ("batch_shape", ctypes.POINTER(ctypes.c_int64)),
        ("batch_strides", ctypes.POINTER(ctypes.c_int64)),
    ]


# Element types of device input (EDA_SRC_* of include/eda_hip.h), by NumPy / PyTorch dtype name.
SOURCE_CODES = {
    "bool": 1, "int8": 2, "int16": 3, "int32": 4, "int64": 5, "uint8": 6, "uint16": 7, "uint32": 8,
    "uint64": 9, "float16": 10, "bfloat16": 11, "float32": 12, "float64": 13,
}
SOURCE_SIZES = {
    "bool": 1, "int8": 1, "int16": 2, "int32": 4, "int64": 8, "uint8": 1, "uint16": 2, "uint32": 4,
    "uint64": 8, "float16": 2, "bfloat16": 2, "float32": 4, "float64": 8,
}


class Params(ctypes.Structure):
    _fields_ = [
        ("max_imf", ctypes.c_int32),
        ("sifts", ctypes.c_int32),
        ("check_finite", ctypes.c_int32),
        ("orientation", ctypes.c_int32),
        ("tie_tolerance", ctypes.c_double),
        ("max_chunk", ctypes.c_int64),
        ("stream", ctypes.c_uint64),
        ("keep_tape", ctypes.c_int32),
        ("dct_mode", ctypes.c_int32),
    ]


class Result(ctypes.Structure):
    _fields_ = [
        ("handle", ctypes.c_void_p),
        ("batch", ctypes.c_int64),
        ("count", ctypes.c_int64),
        ("rows", ctypes.c_int32),
        ("K", ctypes.c_int32),
        ("dtype", ctypes.c_int32),
        ("device", ctypes.c_int32),
        ("chunk", ctypes.c_int64),
        ("n_imfs", ctypes.POINTER(ctypes.c_int32)),
        ("error", ctypes.POINTER(ctypes.c_int32)),
        ("windows", ctypes.POINTER(ctypes.c_int32)),
        ("sifts", ctypes.POINTER(ctypes.c_int32)),
        ("spacing", ctypes.POINTER(ctypes.c_double)),
        ("energy", ctypes.POINTER(ctypes.c_double)),
    ]


# Every symbol include/eda_hip.h declares, with its argument and result types.
_c = ctypes
_vp, _i32, _i64, _f64, _u64 = _c.c_void_p, _c.c_int32, _c.c_int64, _c.c_double, _c.c_uint64
_i32p, _i64p, _f64p, _u64p, _u8p = (_c.POINTER(_c.c_int32), _c.POINTER(_c.c_int64), _c.POINTER(_c.c_double),
                                    _c.POINTER(_c.c_uint64), _c.POINTER(_c.c_uint8))
_err = [_c.c_char_p, _c.c_size_t]
SIGNATURES = {
    "eda_decompose": ([_c.POINTER(ArrayDesc), _c.POINTER(Params), _c.POINTER(Result), *_err], _c.c_int),
    "eda_collect": ([_c.POINTER(Result), _vp, _i32, *_err], _c.c_int),
    "eda_collect_orientation": ([_c.POINTER(Result), _vp, _vp, _i32, *_err], _c.c_int),
    "eda_monogenic": ([_c.POINTER(ArrayDesc), _i32, _i64, _c.c_uint64, _vp, _i32, _i32p, _c.POINTER(_i64), *_err],
                      _c.c_int),
    "eda_release": ([_c.POINTER(Result)], None),
    "eda_backward": ([_c.POINTER(Result), _vp, _vp, _i32, *_err], _c.c_int),
    "eda_replay": ([_c.POINTER(Result), _vp, _vp, *_err], _c.c_int),
    "eda_tape_mode": ([_c.POINTER(Result), _c.c_int64, _c.c_int32, _c.POINTER(_c.c_int32), _c.POINTER(_c.c_double),
                       _vp, _vp, _vp, *_err], _c.c_int),
    "eda_build_info": ([_c.c_char_p, _c.c_size_t], _c.c_int),
    "eda_device_count": ([], _c.c_int),
    "eda_device_info": ([_i32, _c.c_char_p, _c.c_size_t], _c.c_int),
    "eda_pointer_device": ([_vp, _i32p], _c.c_int),
    "eda_timer_enable": ([_i32], _c.c_int),
    "eda_timer_read": ([_f64p, _i64p, _f64p, *_err], _c.c_int),
    "eda_timer_report": ([_c.c_char_p, _c.c_size_t, *_err], _c.c_int),
    "eda_autotune_report": ([_c.c_char_p, _c.c_size_t, *_err], _c.c_int),
    "eda_prim_cube": ([_i32, _i32, _i64, _i64p, _i32, _i32p, _vp, _vp, _vp, *_err], _c.c_int),
    "eda_prim_stats": ([_i32, _i32, _i64, _i64p, _vp, _f64p, _f64p, _u64p, *_err], _c.c_int),
    "eda_prim_select": ([_i32, _i32, _i64, _i64p, _vp, _vp, _i32, _i32, _f64, _f64p, _i64p, *_err], _c.c_int),
    "eda_prim_select_noalloc": ([_i32, _i32, _i64, _i64p, _vp, _vp, _i32, _i32, _f64, _vp, _vp, _u64, *_err],
                                _c.c_int),
    "eda_prim_noise": ([_i32, _i32, _i64, _i64p, _vp, _i32, _f64p, *_err], _c.c_int),
    "eda_prim_tau": ([_i32, _i32, _i64, _i64p, _vp, _i32p, _i32p, _f64p, _f64p, _vp, _i32p, *_err], _c.c_int),
    "eda_prim_flags": ([_i32, _i32, _i64, _i64p, _vp, _vp, _vp, _u64p, *_err], _c.c_int),
    "eda_prim_raw_window": ([_i32, _i32, _i64, _i64p, _vp, _i32p, _vp, *_err], _c.c_int),
    "eda_prim_ccl": ([_i32, _i64, _i64p, _vp, _i32, _vp, *_err], _c.c_int),
    "eda_prim_dct": ([_i32, _i32, _i64, _i64p, _vp, _vp, _i32, _i32, *_err], _c.c_int),
    "eda_prim_orientation": ([_i32, _i32, _i64, _i64p, _vp, _i32p, _vp, _vp, *_err], _c.c_int),
}

_lib = None
_lock = threading.Lock()


def core() -> ctypes.CDLL:
    """The loaded library (raises RuntimeError when it is not built)."""
    global _lib
    if _lib is not None:
        return _lib
    with _lock:
        if _lib is None:
            if not LIB_PATH.exists():
                raise RuntimeError(
                    f"eda: the HIP library {LIB_PATH} is not built; run __graft_entry__.build() "
                    "(hipcc --offload-arch=gfx950). eda has no CPU fallback."
                )
            lib = ctypes.CDLL(str(LIB_PATH))
            for name, (args, res) in SIGNATURES.items():
                fn = getattr(lib, name)
                fn.argtypes = args
                fn.restype = res
            _lib = lib
    return _lib


def check(rc: int, err: ctypes.Array) -> None:
    """Raises the Python exception of a C ABI return code."""
    if rc == EDA_OK:
        return
    from ._errors import CudaError

    msg = err.value.decode(errors="replace") or f"eda: error code {rc}"
    if rc == EDA_EINVAL:
        raise ValueError(msg)
    if rc == EDA_ENOMEM:
        raise MemoryError(msg)
    if rc == EDA_EHIP:
        raise CudaError(msg)
    raise RuntimeError(msg)


def errbuf() -> ctypes.Array:
    return ctypes.create_string_buffer(1024)
