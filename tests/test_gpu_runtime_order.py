"""One HIP runtime per process: creating a server (the library's first GPU
use) before any torch call must leave torch able to use the GPU, and device
memory must be shared between the two."""
import os
import subprocess
import sys

import pytest

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))

PROG = r"""
import numpy as np
from deeprwcap_amd import structures as S
from deeprwcap_amd.serving import Server
sv = Server(0, None, max_batch=64, fdm_tol=1e-6)
L, C, _ = S.layered_boxes(4, 16, seed=3)
out = sv.wait_many(sv.submit_many([0] * 4, [16] * 4, [2] * 4, L, C))
sv.close()
import torch
x = torch.arange(8, dtype=torch.float32, device="cuda")
assert float(x.sum()) == 28.0
assert np.isfinite(out[:, :3]).all()
print("ok", flush=True)
"""


def test_library_first_then_torch():
    r = subprocess.run([sys.executable, "-c", PROG], cwd=ROOT, capture_output=True, text=True, timeout=180)
    print(r.stdout, r.stderr[-2000:], flush=True)
    assert r.returncode == 0 and "ok" in r.stdout
