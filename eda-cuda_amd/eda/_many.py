# SPDX-License-Identifier: Apache-2.0
"""``eda.decompose_many``: a ragged batch, fields of uneven lengths and shapes, in one call.

The reference decomposes one batch of equal fields per call (cpp/python/decompose.cpp:40-136 reads
one array; cpp/src/method/driver.cpp:754-923 runs it in chunks). A caller with signals of many
lengths calls it once per length; most of such a call is the method's per-mode host decisions, so
the GPU idles between them. Here the fields are grouped by shape, dtype and device: every group is
one batched decomposition (its elements share every launch), and the groups run at the same time
from a pool of host threads, each with its own HIP stream (the library keeps one stream per thread
and device), so that one group's launches fill the GPU while another group's host decides.

Each field's rows, information and errors are exactly those of ``eda.decompose`` of that field
alone: every element of a batch is decomposed on its own, and a group's rows beyond an element's
own count are trimmed."""

from __future__ import annotations

import os
from concurrent.futures import ThreadPoolExecutor
from typing import Any, Sequence

import numpy as np

from ._api import (
    ERROR_INTERNAL,
    DecompositionInfo,
    _dct_mode,
    _error_summary,
    _flag,
    _is_device_input,
    _max_imf,
    _sifts,
    _tie_tolerance,
    decompose,
)
from ._errors import DecompositionError


def _tensor(x: Any) -> Any:
    """A device field as a PyTorch tensor (no copy): a tensor, a DLPack import, or an array
    interface wrapped by torch.as_tensor."""
    import torch  # noqa: PLC0415

    if isinstance(x, torch.Tensor):
        return x
    if hasattr(x, "__dlpack__"):
        return torch.from_dlpack(x)
    return torch.as_tensor(x, device="cuda")


def _key(x: Any) -> tuple:
    if _is_device_input(x):
        t = _tensor(x)
        return ("device", t.device.index or 0, str(t.dtype), tuple(t.shape))
    a = np.asarray(x)
    return ("host", -1, a.dtype.str, a.shape)


def _stack(fields: list) -> Any:
    if _is_device_input(fields[0]):
        import torch  # noqa: PLC0415

        return torch.stack([_tensor(f) for f in fields])
    return np.stack([np.asarray(f) for f in fields])


def _item_info(info: DecompositionInfo, b: int, ndim: int) -> DecompositionInfo:
    k = int(info.n_imfs[b])
    return DecompositionInfo(
        n_imfs=np.asarray(info.n_imfs[b]),
        error=np.asarray(info.error[b]),
        windows=info.windows[b, :k].copy(),
        spacing=info.spacing[b, :k].copy(),
        sifts=info.sifts[b, :k].copy(),
        energy=info.energy[b, :k].copy(),
        dtype=info.dtype,
        axes=tuple(range(ndim)),
    )


def decompose_many(
    fields: Sequence[Any],
    max_imf: int = -1,
    *,
    dtype: Any = None,
    tie_tolerance: float | None = None,
    sifts: int = 3,
    check_finite: bool = True,
    errors: str = "raise",
    return_info: bool = False,
    dct: str = "exact",
    workers: int | None = None,
) -> list:
    """Decompose every field of a ragged batch: ``fields`` is a sequence of 1-D, 2-D or 3-D
    arrays (host arrays or ROCm device tensors) of any shapes. Returns a list, one entry per
    field in order: its rows (P, *shape) as ``eda.decompose(field)`` returns them, or (rows,
    info) with ``return_info``. Fields of one shape, dtype and device run as one batch; the
    batches run concurrently on ``workers`` host threads (default: min(8, batches, CPUs)), each
    with its own stream. ``errors="raise"`` raises one ``DecompositionError`` for the failed
    fields (``indices`` are positions in ``fields``) after every batch ends; ``"status"``
    returns NaN rows and the codes in each field's info, as ``eda.decompose`` does."""
    max_imf = _max_imf(max_imf)
    tie = _tie_tolerance(tie_tolerance)
    sifts = _sifts(sifts)
    check_finite = _flag(check_finite, "check_finite")
    return_info = _flag(return_info, "return_info")
    _dct_mode(dct)
    if not isinstance(errors, str) or errors not in ("raise", "status"):
        raise ValueError(f"eda: errors={errors!r} is not one of 'raise', 'status'")
    items = list(fields)
    groups: dict[tuple, list[int]] = {}
    for i, f in enumerate(items):
        groups.setdefault(_key(f), []).append(i)
    for key in groups:
        if not 1 <= len(key[3]) <= 3:
            raise ValueError(f"eda: every field needs 1 to 3 axes, got shape {key[3]}")
    # the largest batches first, so that the pool ends on small ones
    order = sorted(groups.values(), key=lambda idx: -len(idx) * int(np.prod(_key(items[idx[0]])[3])))
    if workers is None:
        workers = min(8, len(order), os.cpu_count() or 1)
    workers = max(1, int(workers))

    def run(idx: list[int]):
        batch = _stack([items[i] for i in idx])
        y, info = decompose(batch, axes=tuple(range(1, batch.ndim)), max_imf=max_imf, dtype=dtype,
                            tie_tolerance=tie, sifts=sifts, check_finite=check_finite, errors="status",
                            return_info=True, dct=dct)
        return idx, y, info

    out: list = [None] * len(items)
    codes = np.zeros(len(items), np.int32)
    if workers == 1 or len(order) == 1:
        done = [run(idx) for idx in order]
    else:
        with ThreadPoolExecutor(max_workers=workers) as ex:
            done = list(ex.map(run, order))
    for idx, y, info in done:
        for b, i in enumerate(idx):
            err = int(info.error[b])
            codes[i] = err
            P = int(info.n_imfs[b]) + 1 if err == 0 else 1
            rows = y[b, :P]
            out[i] = (rows, _item_info(info, b, y.ndim - 2)) if return_info else rows
    internal = np.flatnonzero(codes == ERROR_INTERNAL)
    if internal.size:
        raise RuntimeError(
            "eda: an internal invariant failed (error code 255) for field "
            f"{', '.join(str(int(i)) for i in internal[:8])}: this is a bug in eda"
        )
    failed = np.flatnonzero(codes != 0)
    if failed.size and errors == "raise":
        raise DecompositionError(_error_summary(codes[failed], failed, codes.size), indices=failed,
                                 codes=codes[failed], info=None)
    return out
