# SPDX-License-Identifier: Apache-2.0
"""The argument rules of eda.decompose (CPU: every rule raises before any GPU work), restated from
the reference's tests/python/test_api.py:100-210 and src/eda/_api.py:433-526, and the layout
permutation of the result."""

from __future__ import annotations

import numpy as np
import pytest

import eda
from eda._api import _normalize_axes, _restore_order

X = np.sin(np.arange(64.0))


@pytest.mark.parametrize("bad", [True, 1.5, "3", None])
def test_max_imf_must_be_an_integer(bad):
    with pytest.raises(TypeError):
        eda.decompose(X, max_imf=bad)


@pytest.mark.parametrize("bad,exc", [(0, ValueError), (-1, ValueError), (True, TypeError), (2.0, TypeError)])
def test_sifts_rules(bad, exc):
    with pytest.raises(exc):
        eda.decompose(X, sifts=bad)


@pytest.mark.parametrize("bad,exc", [(-1.0, ValueError), (float("inf"), ValueError), (float("nan"), ValueError),
                                     (True, TypeError), ("1e-3", TypeError)])
def test_tie_tolerance_rules(bad, exc):
    with pytest.raises(exc):
        eda.decompose(X, tie_tolerance=bad)


def test_flags_must_be_bool():
    for name in ("check_finite", "return_info", "return_orientation"):
        with pytest.raises(TypeError):
            eda.decompose(X, **{name: 1})


def test_errors_policy_name():
    with pytest.raises(ValueError):
        eda.decompose(X, errors="ignore")


def test_dtype_rules():
    with pytest.raises(TypeError):
        eda.decompose(np.arange(64))  # int64 without dtype=
    with pytest.raises(TypeError):
        eda.decompose(X, dtype="float16")
    with pytest.raises(TypeError):
        eda.decompose(X.astype(np.complex128))


def test_axes_rules():
    assert _normalize_axes(None, 2) == (0, 1)
    assert _normalize_axes(-1, 3) == (2,)
    assert _normalize_axes((2, 0), 3) == (2, 0)
    with pytest.raises(ValueError):
        _normalize_axes(None, 4)
    with pytest.raises(ValueError):
        _normalize_axes((0, 0), 2)
    with pytest.raises(np.exceptions.AxisError):
        _normalize_axes(3, 2)
    with pytest.raises(ValueError):
        _normalize_axes((0, 1, 2, 3), 4)
    with pytest.raises(TypeError):
        _normalize_axes((True,), 2)


def test_return_orientation_needs_2_axes():
    with pytest.raises(ValueError, match="return_orientation=True needs 2 or 3 spatial axes"):
        eda.decompose(X, return_orientation=True)
    with pytest.raises(ValueError, match="needs 2 or 3 spatial axes, got 1"):
        eda.decompose(np.zeros((4, 64)), axes=1, return_orientation=True)
    with pytest.raises(TypeError, match="return_orientation must be True or False"):
        eda.decompose(np.zeros((16, 16)), return_orientation=1)
    assert eda.Orientation.__dataclass_fields__.keys() == {"direction", "coherence"}


def test_restore_order_of_the_orientation():
    # direction (B, K, D, *spatial) of x (b0, s0, b1, s1), axes (1, 3) -> (b0, K, D, s0, b1, s1)
    perm = _restore_order(4, (0, 2), (1, 3), 2)
    assert np.zeros((2, 3, 7, 2, 5, 6)).transpose(perm).shape == (2, 7, 2, 5, 3, 6)


def test_restore_order_inserts_the_mode_axis_before_the_first_spatial_axis():
    # x of shape (b0, s0, b1, s1), axes (1, 3): computed (b0, b1, P, s0, s1) -> (b0, P, s0, b1, s1)
    perm = _restore_order(4, (0, 2), (1, 3))
    comp = np.zeros((2, 3, 7, 5, 6))
    assert comp.transpose(perm).shape == (2, 7, 5, 3, 6)


# ----------------------------------------------------------------------------------------------
# The reference's native-module interface (src/eda/_core.pyi:9-27) served over the C ABI

_CORE_SIGNATURES = {
    "build_info": [],
    "device_count": [],
    "device_info": ["device"],
    "pointer_device": ["ptr"],
    "decompose_host": ["x", "max_imf", "tie_tolerance", "sifts", "check_finite", "orientation", "max_chunk"],
    "monogenic_host": ["x", "check_finite", "max_chunk"],
    "monogenic_device": ["ptr", "device", "dtype", "target", "batch_shape", "batch_strides", "shape", "strides",
                         "fences", "check_finite", "max_chunk"],
    "decompose_device": ["ptr", "device", "dtype", "target", "batch_shape", "batch_strides", "shape", "strides",
                         "fences", "max_imf", "tie_tolerance", "sifts", "check_finite", "orientation", "max_chunk"],
}


@pytest.mark.parametrize("name", sorted(_CORE_SIGNATURES))
def test_core_hip_has_the_reference_interface(name):
    import inspect

    from eda import _core_hip

    fn = getattr(_core_hip, name)
    assert list(inspect.signature(fn).parameters) == _CORE_SIGNATURES[name]


def test_core_hip_argument_rules_before_any_gpu_work():
    from eda import _core_hip

    x = np.zeros((2, 16), np.float32)
    with pytest.raises(ValueError):
        _core_hip.decompose_host(x, -1, None, 0, True, False, 0)  # sifts
    with pytest.raises(TypeError):
        _core_hip.decompose_host(x.astype(np.int32), -1, None, 3, True, False, 0)
    with pytest.raises(ValueError):
        _core_hip.decompose_host(np.zeros((2, 4, 4), np.float32)[:, ::2], -1, None, 3, True, False, 0)


@pytest.mark.parametrize("shape,strides,want", [
    ((4,), (100,), ([4], [100])),
    ((2, 3), (300, 100), ([6], [100])),
    ((2, 1, 3), (300, 7, 100), ([6], [100])),
    ((), (), ([], [])),
    ((3, 2), (-200, -100), ([6], [-100])),
    ((2, 3), (100, 300), ([2, 3], [100, 300])),  # does not merge: two axes
    ((1, 1), (5, 9), ([], [])),
])
def test_core_hip_merged_batch_axes(shape, strides, want):
    """The reference's merge of the batch axes (cpp/src/decompose.cpp:48-59): extent-1 axes dropped,
    neighbours whose strides chain merged."""
    from eda._core_hip import _merged_batch_axes

    assert _merged_batch_axes(shape, strides) == want


def test_core_hip_device_descriptor_rules():
    from eda import _native
    from eda._core_hip import _device_desc

    desc, keep = _device_desc(4096, 0, "int16", "float32", (2, 3), (100, 300), (5, 7), (7, 1))
    assert desc.source == _native.SOURCE_CODES["int16"] and desc.batch_axes == 2 and desc.batch == 6
    assert [desc.batch_shape[k] for k in range(2)] == [2, 3] and [desc.batch_strides[k] for k in range(2)] == [100, 300]
    desc, keep = _device_desc(4096, 0, "float32", "float32", (4,), (35,), (5, 7), (7, 1))
    assert desc.source == 0 and desc.batch_axes == 0 and desc.batch_stride == 35
    with pytest.raises(TypeError):
        _device_desc(4096, 0, "complex64", "float32", (4,), (35,), (5, 7), (7, 1))
    with pytest.raises(ValueError):
        _device_desc(4096, 0, "float32", "float16", (4,), (35,), (5, 7), (7, 1))


@pytest.mark.parametrize(("kwargs", "error", "fragment"), [
    (dict(check_finite=1), TypeError, "eda: check_finite must be True or False, got int"),
    (dict(dtype="float16"), TypeError, "eda: dtype= must be float32 or float64, got float16"),
    (dict(axes=(0, 0)), ValueError, "eda: axes must be unique"),
    (dict(axes=()), ValueError, "eda: 1 to 3 spatial axes, got 0"),
])
def test_monogenic_argument_errors(kwargs, error, fragment):
    import re

    with pytest.raises(error, match=re.escape(fragment)):
        eda.monogenic(np.zeros((4, 16), np.float32), **kwargs)


def test_monogenic_result_type():
    assert set(eda.Monogenic.__dataclass_fields__) == {"amplitude", "phase", "frequency", "direction", "coherence"}
    with pytest.raises(TypeError, match="expected a float32 or float64 array"):
        eda.monogenic(np.zeros((4, 16), np.int32))


def test_dct_mode_rules():
    from eda._api import _dct_mode

    assert _dct_mode("exact") == 0 and _dct_mode("bf16x3") == 1
    for bad in ("fp32", None, 1):
        with pytest.raises(ValueError):
            _dct_mode(bad)
