# SPDX-License-Identifier: Apache-2.0
"""pytest setup: the package (eda-cuda_amd/eda) and the oracle (oracle/) on sys.path, the `gpu`
marker (tests that need an AMD GPU and the built HIP library)."""

import sys
from pathlib import Path

ROOT = Path(__file__).resolve().parents[1]
for p in (ROOT / "eda-cuda_amd", ROOT / "oracle", ROOT / "tests"):
    if str(p) not in sys.path:
        sys.path.insert(0, str(p))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an AMD GPU (gfx950) and the built libeda_hip.so")
