"""The library loaded before PyTorch has initialised its HIP runtime: a fresh
process imports the package and builds a mesh (which loads the library),
then initialises torch on the GPU, then creates a device context and runs a
residual.  torch ships its own HIP runtime; with the library's loaded first
the library's contexts found no device, so the package loads torch first.
One subprocess, so that nothing in this test process has touched the GPU
or the library beforehand."""
import os
import subprocess
import sys

import pytest

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))

SCRIPT = r"""
import sys
sys.path.insert(0, sys.argv[1])
import numpy as np
import intpointsfem_amd as ipf
from intpointsfem_amd import wellbore as wb
m = wb.make_mesh((6, 5, 2, 1.1))
import torch
assert torch.cuda.is_available()
x = torch.ones(8, device="cuda")
sm = ipf.StructMatrix(m, device=0)
ctx = sm.context()
neq = sm.sizes()["neq"]
u = torch.as_tensor(wb.lame_displacement(m.nodes(), wb.normal_stress(10)), device="cuda")
r = torch.zeros(neq, dtype=torch.float64, device="cuda")
ctx.residual(u, r)
torch.cuda.synchronize()
print("ok", neq, float(x.sum()), bool(torch.isfinite(r).all()), flush=True)
"""


def test_library_loaded_before_torch(gpu):
    out = subprocess.run([sys.executable, "-c", SCRIPT, ROOT], capture_output=True, text=True, timeout=300)
    print(out.stdout[-2000:], out.stderr[-2000:])
    assert out.returncode == 0
    assert "ok 286 8.0 True" in out.stdout
