"""The library and PyTorch share one HIP runtime whatever is used first.

PyTorch's HIP libraries load their own libamdhip64.so; libbine_hip.so loaded
before them would bring the system runtime as a second one, and PyTorch's
would then find no GPU.  A fresh interpreter uses the library (a link
simulation) before touching torch, then allocates a torch tensor on the GPU.
"""
import subprocess
import sys
import os

import pytest

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))

SCRIPT = r"""
import sys
sys.path.insert(0, %r)
from bine_trees_fugaku_amd import linksim as ls
sim = ls.LinkSim([([4, 4], "allreduce", "bine")])
sim.run()
st = sim.stats(0)
sim.close()
import torch
x = torch.ones(4, device="cuda")
assert float(x.sum()) == 4.0
print("one runtime: ok", flush=True)
"""


def test_library_before_torch(gpu):
    r = subprocess.run([sys.executable, "-c", SCRIPT % ROOT], capture_output=True, text=True, timeout=240)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "one runtime: ok" in r.stdout
