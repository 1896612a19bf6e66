# SPDX-License-Identifier: Apache-2.0
"""The exceptions of eda (importable without the compiled library); both pickle. Same classes and
meaning as the reference's src/eda/_errors.py:10-42."""

from __future__ import annotations

from typing import Any


class CudaError(RuntimeError):
    """A HIP runtime call or kernel launch failed inside eda, or no usable GPU was found.

    The name follows the reference's API (eda.CudaError) so that callers' handlers keep working;
    ``code`` and ``name`` are the HIP error's number and symbolic name when known."""

    def __init__(self, message: str, code: int | None = None, name: str | None = None) -> None:
        super().__init__(message)
        self.code = code
        self.name = name or None

    def __reduce__(self):
        return (type(self), (str(self), self.code, self.name))


class DecompositionError(ValueError):
    """Raised by ``eda.decompose(..., errors="raise")`` when one or more batch elements could not be
    decomposed; the elements that succeeded are not returned.

    Attributes: ``indices``, the failed elements' positions in the batch counted in C order;
    ``codes``, one error code per failed element (1 when ``check_finite`` found inf or NaN in its
    input, 2 when an intermediate field or an output row became non-finite); ``info``, the call's
    ``DecompositionInfo``, whose ``error`` array holds the code of every element."""

    def __init__(self, message: str, indices: Any = (), codes: Any = (), info: Any = None) -> None:
        super().__init__(message)
        self.indices = indices
        self.codes = codes
        self.info = info

    def __reduce__(self):
        return (type(self), (str(self), self.indices, self.codes, self.info))
