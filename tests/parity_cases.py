# SPDX-License-Identifier: Apache-2.0
"""The reference's parity manifest as test data: the cases and their decisions
(tests/golden/parity_decisions.json), the inputs (tests/golden/parity_inputs.npz, written by
tests/golden/make_parity_inputs.py and checked against input_sha256 by the CPU tests), and the
helpers that compare decisions."""

from __future__ import annotations

import json
from functools import lru_cache
from pathlib import Path

import numpy as np

GOLDEN = Path(__file__).resolve().parent / "golden"
MANIFEST = json.loads((GOLDEN / "parity_decisions.json").read_text())
CASES = MANIFEST["cases"]


@lru_cache(maxsize=1)
def _inputs():
    return np.load(GOLDEN / "parity_inputs.npz", allow_pickle=False)


def load_input(case_id: str) -> np.ndarray:
    """The input array of a case (its dtype, shape and layout)."""
    return np.asarray(_inputs()[case_id])


def batch_layout(x: np.ndarray, axes) -> tuple[np.ndarray, tuple[int, ...]]:
    """x as (B, *spatial) with the spatial axes in the order of axes and the batch in C order."""
    axes = tuple(axes) if axes is not None else tuple(range(x.ndim))
    batch = [i for i in range(x.ndim) if i not in axes]
    spatial = tuple(x.shape[a] for a in axes)
    return np.ascontiguousarray(x.transpose(batch + list(axes))).reshape((-1,) + spatial), axes


def decisions_equal(c: dict, n_imfs, error, windows, sifts) -> list[str]:
    """The decision fields that differ from the case's (empty when all equal)."""
    bad = []
    if list(np.asarray(n_imfs).reshape(-1)) != c["n_imfs"]:
        bad.append("n_imfs")
    if list(np.asarray(error).reshape(-1)) != c["error"]:
        bad.append("error")
    if c["windows"] is not None and list(np.asarray(windows).reshape(-1)) != c["windows"]:
        bad.append("windows")
    if c["sifts"] is not None and list(np.asarray(sifts).reshape(-1)) != c["sifts"]:
        bad.append("sifts")
    return bad
