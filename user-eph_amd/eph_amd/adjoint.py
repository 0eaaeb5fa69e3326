"""torch.autograd binding of the trajectory adjoint (include/eph_adjoint.h).

    traj = Trajectory(dict(cells=20, dt=5e-4, nsteps=200, every=20, pka_eV=2000.0,
                           T=600.0, target_seed=1))
    beta = torch.tensor(traj.beta0, requires_grad=True)
    v0 = torch.tensor(traj.v0, requires_grad=True)      # optional
    loss = trajectory_loss(beta, v0, traj)
    loss.backward()                                      # beta.grad, v0.grad

The GPU work runs in a child process (`python -m eph_amd.adjoint_worker`) that
loads libeph_amd.so and never imports torch; PyTorch wheels bundle a second
HIP runtime, and one process must not drive the GPU through both.  Tensors
cross the pipe as pickled CPU float64 arrays.
"""
import os
import pickle
import subprocess
import sys

import numpy as np
import torch

_PKG_PARENT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


class Trajectory:
    """A trajectory held by a worker process; spec as adjoint_worker.build_trajectory."""

    def __init__(self, spec):
        env = dict(os.environ)
        env["PYTHONPATH"] = _PKG_PARENT + os.pathsep + env.get("PYTHONPATH", "")
        self._proc = subprocess.Popen([sys.executable, "-m", "eph_amd.adjoint_worker"],
                                      stdin=subprocess.PIPE, stdout=subprocess.PIPE, env=env)
        self._send(spec)
        reply = self._recv()
        if reply[0] != "ok":
            self.close()
            raise RuntimeError(f"adjoint worker: {reply[1]}")
        info = reply[1]
        self.n_elem, self.n_beta = info["n_elem"], info["n_beta"]
        self.nlocal, self.ncells = info["nlocal"], info["ncells"]
        self.beta0 = info["beta0"]
        self.v0 = info["v0"]

    def _send(self, obj):
        pickle.dump(obj, self._proc.stdin, protocol=pickle.HIGHEST_PROTOCOL)
        self._proc.stdin.flush()

    def _recv(self):
        try:
            return pickle.load(self._proc.stdout)
        except EOFError:
            raise RuntimeError(f"adjoint worker ended (status {self._proc.wait()})")

    def _call(self, *msg):
        self._send(msg)
        reply = self._recv()
        if reply[0] != "ok":
            raise RuntimeError(f"adjoint worker: {reply[1]}")
        return reply[1:]

    def forward(self, beta=None, v0=None):
        return self._call("forward", beta, v0)

    def backward(self):
        return self._call("backward")

    def close(self):
        if getattr(self, "_proc", None) is not None:
            try:
                self._send(("close",))
                self._proc.stdin.close()
            except (BrokenPipeError, OSError, ValueError):
                pass
            try:
                self._proc.wait(60)
            except subprocess.TimeoutExpired:
                self._proc.kill()
                self._proc.wait()
            self._proc = None

    def __del__(self):
        self.close()


class _TrajectoryLoss(torch.autograd.Function):
    @staticmethod
    def forward(ctx, beta, v0, traj):
        b = beta.detach().cpu().numpy().astype(np.float64)
        v = None if v0 is None else v0.detach().cpu().numpy().astype(np.float64)
        loss, _, _ = traj.forward(b, v)
        ctx.traj = traj
        ctx.has_v0 = v0 is not None
        return torch.tensor(loss, dtype=torch.float64)

    @staticmethod
    def backward(ctx, grad_out):
        gb, gv = ctx.traj.backward()
        g = grad_out.item()
        gbeta = torch.from_numpy(gb * g)
        gv0 = torch.from_numpy(gv * g) if ctx.has_v0 else None
        return gbeta, gv0, None


def trajectory_loss(beta, v0, traj):
    """L(beta knots [n_elem, n_beta], initial velocities [nlocal, 3] or None)"""
    return _TrajectoryLoss.apply(beta, v0, traj)
