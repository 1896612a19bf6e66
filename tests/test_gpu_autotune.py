# SPDX-License-Identifier: Apache-2.0
"""The launch-shape autotuner (GPU): a decomposition with the tuned shapes (the default) equals,
bit for bit, the one with the rule's shapes (EDA_AUTOTUNE=0, in a child process); the report names
the tuned kernels (the streaming strided pass's piece count, the pipelined box sum's workgroups per
CU) with a choice inside each candidate list."""

from __future__ import annotations

import ctypes
import json
import os
import subprocess
import sys
from pathlib import Path

import numpy as np
import pytest

import eda
import workloads

pytestmark = pytest.mark.gpu
ROOT = Path(__file__).resolve().parents[1]

CHILD = r"""
import sys
sys.path[:0] = [sys.argv[1] + '/eda-cuda_amd', sys.argv[1] + '/tests']
import numpy as np, torch, eda, workloads
x = torch.from_numpy(workloads.rings_stripes_blobs(6, 256)).to('cuda')
y = eda.decompose(x, axes=(1, 2))
np.save(sys.argv[2], y.cpu().numpy())
"""


def report():
    from eda import _native

    lib = _native.core()
    buf = ctypes.create_string_buffer(1 << 20)
    err = _native.errbuf()
    _native.check(lib.eda_autotune_report(buf, len(buf), err, len(err)), err)
    return json.loads(buf.value.decode())


def test_tuned_shapes_give_the_rule_shapes_bits(tmp_path):
    import torch

    x = torch.from_numpy(workloads.rings_stripes_blobs(6, 256)).to("cuda")
    y = eda.decompose(x, axes=(1, 2)).cpu().numpy()
    out = tmp_path / "rule.npy"
    env = dict(os.environ, EDA_AUTOTUNE="0")
    print("autotune: child process with EDA_AUTOTUNE=0", flush=True)
    subprocess.run([sys.executable, "-c", CHILD, str(ROOT), str(out)], check=True, env=env, timeout=300)
    y0 = np.load(out)
    assert np.array_equal(y.view(np.uint8), y0.view(np.uint8))
    rep = report()
    kinds = {k.split(":", 1)[1].split(":", 1)[0] for k in rep}
    assert "k_vhgw_stream" in kinds and "k_box" in kinds, sorted(rep)[:5]
    for key, v in rep.items():
        assert 0 <= v["choice"] < len(v["us"]), key
        assert all(t > 0 for t in v["us"]), key
