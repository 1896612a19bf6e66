# SPDX-License-Identifier: Apache-2.0
"""The C ABI (CPU): the library loads, exports every entry point include/eda_hip.h declares, and
fails loudly (an exception, never a CPU fallback) when no GPU is visible."""

from __future__ import annotations

import ctypes
import re
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]
HEADER = ROOT / "include" / "eda_hip.h"


def declared_functions() -> list[str]:
    text = HEADER.read_text()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(eda_[a-z_]+)\s*\(", text)))


def test_header_declares_the_boundary():
    names = declared_functions()
    for required in ("eda_decompose", "eda_collect", "eda_release", "eda_device_count", "eda_build_info"):
        assert required in names


def test_library_exports_every_declared_symbol():
    from eda import _native

    lib = ctypes.CDLL(str(_native.LIB_PATH))
    for name in declared_functions():
        assert hasattr(lib, name), name
    assert set(_native.SIGNATURES) == set(declared_functions())


def test_build_info_without_a_gpu():
    import eda

    info = eda.build_info()
    assert info["architectures"] == ["gfx950"] and info["backend"] == "hip"


def test_decompose_raises_instead_of_falling_back():
    import numpy as np

    import eda

    if eda.device_count() > 0:
        pytest.skip("a GPU is visible")
    with pytest.raises(eda.CudaError):
        eda.decompose(np.sin(np.arange(64.0)))
