"""Runs the torch.autograd.Function check harness (tests/adjoint_harness.py) in its own
process: the GPU splat / sampling adjoints at the 2,097,152-point size against torch autograd
over a dense FP64 grid."""
import json
import os
import subprocess
import sys

import pytest

import gpu_util as U

HARNESS = os.path.join(os.path.dirname(os.path.abspath(__file__)), "adjoint_harness.py")


@pytest.mark.gpu
@pytest.mark.timeout(600)
def test_autograd_harness_against_dense_reference():
    U.log("case autograd harness (child process)")
    r = subprocess.run([sys.executable, "-u", HARNESS], capture_output=True, text=True, timeout=560)
    U.log(r.stdout[-3000:])
    assert r.returncode == 0, r.stdout[-3000:] + r.stderr[-3000:]
    res = json.loads([l for l in r.stdout.splitlines() if l.startswith("{")][-1])
    for k, v in res.items():
        if k.endswith("dense"):
            assert v <= 1e-5, (k, v)
