"""The binding loads PyTorch's HIP runtime before the library: a process that loads
the library before importing torch still has one runtime, a visible GPU, and a
filter that steps on torch tensors (run in a child process, so the order is real)."""
import os
import subprocess
import sys

import pytest

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "newton-stein-variational-particle-filter_amd")

CHILD = r"""
import sys
sys.path.insert(0, sys.argv[1])
from svpf_amd import _lib
_lib.load()                      # the library first
import torch                     # then torch
assert torch.cuda.is_available(), "no GPU after loading the library first"
from svpf_amd import config
from svpf_amd.filter import Filter
f = Filter(config.plain_config(n=256))
f.initialize(7)
y = torch.tensor([0.01, -0.02, 0.015, 0.03], device="cuda")
f.run_device(y)
torch.cuda.synchronize()
vol = float(f.last_outputs()[1])
assert vol == vol and vol > 0.0, vol
print("ok", vol)
f.close()
"""


def test_library_loaded_before_torch(torch_cuda):
    r = subprocess.run([sys.executable, "-c", CHILD, PKG], capture_output=True, text=True, timeout=240)
    assert r.returncode == 0, r.stderr[-2000:]
    assert r.stdout.strip().startswith("ok")
