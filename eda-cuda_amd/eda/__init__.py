# SPDX-License-Identifier: Apache-2.0
"""eda on AMD Instinct MI355X: N-dimensional empirical decomposition with hand-written gfx950 HIP
kernels.

``eda.decompose(x)`` splits every 1-D, 2-D or 3-D field of an array into intrinsic mode functions
and a residue; it is a drop-in for the reference package's ``eda.decompose`` (same arguments,
result layout, information and errors). Importing the package needs neither a GPU nor the built
library; the library loads on the first call that needs it, and there is no CPU fallback."""

from __future__ import annotations

import json
import operator
from typing import Any

from . import _native
from ._api import DecompositionInfo, Monogenic, Orientation, decompose, monogenic
from ._autograd import TapedDecomposition, decompose_function, differentiable_decompose
from ._device import DeviceArray
from ._errors import CudaError, DecompositionError
from ._many import decompose_many

__version__ = "0.1.0"

__all__ = [
    "CudaError",
    "DecompositionError",
    "DecompositionInfo",
    "DeviceArray",
    "Monogenic",
    "Orientation",
    "TapedDecomposition",
    "__version__",
    "build_info",
    "decompose",
    "decompose_function",
    "decompose_many",
    "device_count",
    "differentiable_decompose",
    "device_info",
    "monogenic",
]


def build_info() -> dict[str, Any]:
    """How the library was built (version, HIP version, architectures, compiler); works without
    a GPU."""
    import ctypes

    buf = ctypes.create_string_buffer(4096)
    _native.core().eda_build_info(buf, len(buf))
    return json.loads(buf.value.decode())


def device_count() -> int:
    """The number of visible HIP devices; 0 without a driver or device."""
    return int(_native.core().eda_device_count())


def device_info(device: int = 0) -> dict[str, Any]:
    """Properties of a visible HIP device: name, architecture, compute units, memory."""
    import ctypes

    if isinstance(device, bool):
        raise TypeError("eda: device must be an integer, not bool")
    try:
        number = operator.index(device)
    except TypeError:
        raise TypeError(f"eda: device must be an integer, got {type(device).__name__}") from None
    if not 0 <= number < max(device_count(), 0) or number >= 2**31:
        raise ValueError(f"eda: HIP device {number} is not visible")
    buf = ctypes.create_string_buffer(4096)
    err = _native.core().eda_device_info(number, buf, len(buf))
    if err:
        raise CudaError(buf.value.decode())
    return json.loads(buf.value.decode())
