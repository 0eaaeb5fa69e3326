"""GPU: the torch.autograd binding of the trajectory adjoint (eph_amd.adjoint).

Runs first among the GPU tests and in child processes only: the torch process
never loads libeph_amd.so, and its worker process never imports torch, so no
process drives the GPU through two HIP runtimes.
"""
import json
import os
import subprocess
import sys

import pytest

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SMALL = dict(cells=4, dt=5e-4, nsteps=6, every=4, T=600.0, pka_eV=100.0, target_seed=1,
             n_beta=64, md_seed=7, fix_seed=4711)
H = 1e-6


def test_torch_autograd_function(tmp_path):
    """trajectory_loss in a torch process (the GPU work in its worker process)"""
    script = tmp_path / "torch_adj.py"
    script.write_text(f"""
import json, sys
import numpy as np
sys.path.insert(0, {os.path.join(ROOT, 'user-eph_amd')!r})
import torch
from eph_amd.adjoint import Trajectory, trajectory_loss
spec = {dict(SMALL, tmpdir=str(tmp_path))!r}
traj = Trajectory(spec)
beta = torch.tensor(traj.beta0, requires_grad=True)
v0 = torch.tensor(traj.v0, requires_grad=True)
loss = trajectory_loss(beta, v0, traj)
loss.backward()
rng = np.random.default_rng(5)
d = rng.standard_normal(traj.beta0.shape) * traj.beta0
h = {H!r}
fd = (traj.forward(traj.beta0 + h * d, traj.v0)[0] - traj.forward(traj.beta0 - h * d, traj.v0)[0]) / (2 * h)
out = dict(loss=loss.item(), ad=float((beta.grad.numpy() * d).sum()), fd=fd,
           vgrad_norm=float(v0.grad.norm()))
traj.close()
print(json.dumps(out))
""")
    r = subprocess.run([sys.executable, str(script)], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr[-2000:]
    out = json.loads(r.stdout.strip().splitlines()[-1])
    assert abs(out["fd"] - out["ad"]) <= 1e-5 * abs(out["ad"]) + 1e-9 * abs(out["loss"]), out
    assert out["vgrad_norm"] > 0
