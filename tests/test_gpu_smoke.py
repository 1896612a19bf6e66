# SPDX-License-Identifier: Apache-2.0
"""The package's smoke check (one small decomposition of a 1-D signal and of two 2-D images on the
device, against the oracle) runs as a test, so that an interface change cannot leave it behind."""

from __future__ import annotations

import sys
from pathlib import Path

import pytest

pytestmark = pytest.mark.gpu


def test_smoke_runs():
    root = Path(__file__).resolve().parents[1]
    if str(root) not in sys.path:
        sys.path.insert(0, str(root))
    import __graft_entry__

    __graft_entry__.smoke()
