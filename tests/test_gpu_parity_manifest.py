# SPDX-License-Identifier: Apache-2.0
"""The reference's parity manifest on the GPU (tests/golden/parity_decisions.json, from the
reference's tests/data/parity_v2.json; inputs tests/golden/parity_inputs.npz): every one of the 287
cases through ``eda.decompose`` (and,
where the case sets max_chunk, through the reference's native interface with that chunk), in its
dtype, layout and arguments. The decisions (n_imfs, error, windows, sifts), the output shape, the
chunk count and the error policy equal the manifest's exactly. The manifest's row hashes are the
reference's CUDA bits and are not compared; the rows are checked against the oracle elsewhere.
"""

from __future__ import annotations

import json
import math
from pathlib import Path

import numpy as np
import pytest

import eda
from parity_cases import CASES, batch_layout, decisions_equal, load_input

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("case", CASES, ids=lambda c: c["id"])
def test_gpu_decisions_match_the_manifest(case):
    print("parity case", case["id"], flush=True)
    x = load_input(case["id"])
    axes = tuple(case["axes"]) if case["axes"] is not None else None
    y, info = eda.decompose(x, axes=axes, max_imf=case["max_imf"], dtype=case["dtype"],
                            tie_tolerance=case["tie_tolerance"], check_finite=case["check_finite"],
                            errors=case["errors"], return_info=True)
    assert list(y.shape) == case["y_shape"]
    assert not decisions_equal(case, info.n_imfs, info.error, info.windows, info.sifts)
    if case["max_chunk"] is not None:
        from eda import _core_hip

        xb, ax = batch_layout(x, axes)
        _, raw = _core_hip.decompose_host(xb, case["max_imf"], case["tie_tolerance"], 3, case["check_finite"],
                                          False, case["max_chunk"])
        assert math.ceil(xb.shape[0] / raw["chunk"]) == case["chunks"]
        assert not decisions_equal(case, raw["n_imfs"], raw["error"], raw["windows"], raw["sifts"])
