"""ODEFunc / NeuralODE: a restatement of src/python/layers/neuralode_fast.py:7-64.

* ``ODEFunc``: velocity f(t, y) = MLP(cat(y, t) - 0.5), 4 -> 50 -> 50 -> 50 -> 3
  with LeakyReLU (slope 0.01), weights N(0, 0.1), zero biases, built in the
  reference's module order so that one torch seed gives the same weights.
* ``odeint``: the fixed-grid ``method="rk4"`` of torchdiffeq (a third-party
  package that the reference imports and that is not installed here; its
  version is not pinned): one 3/8-rule step per grid interval,
      k1 = f(t0, y0)
      k2 = f(t0 + dt/3, y0 + dt k1 / 3)
      k3 = f(t0 + 2dt/3, y0 + dt (k2 - k1/3))
      k4 = f(t1, y0 + dt (k1 - k2 + k3))
      y1 = y0 + (k1 + 3 (k2 + k3) + k4) dt / 8,
  where a decreasing grid is integrated as t -> -t with f -> -f(-t, y), as
  torchdiffeq does.  rtol/atol do not apply to a fixed-grid method.
  ``forward`` integrates over [0, 1] and ``inverse`` over [1, 0]: one step
  each.  A grid on another device than y0 is moved to y0's device.
  Gradients are plain autograd through the step, as with torchdiffeq's
  non-adjoint ``odeint``.
"""
from __future__ import annotations

import numpy as np
import torch
from torch import nn

_one_third = 1.0 / 3.0
_two_thirds = 2.0 / 3.0


class ODEFunc(nn.Module):
    def __init__(self):
        super().__init__()
        m = 50
        nlin = nn.LeakyReLU()
        self.net = nn.Sequential(
            nn.Linear(4, m),
            nlin,
            nn.Linear(m, m),
            nlin,
            nn.Linear(m, m),
            nlin,
            nn.Linear(m, 3),
        )
        for mod in self.net.modules():
            if isinstance(mod, nn.Linear):
                nn.init.normal_(mod.weight, mean=0, std=1e-1)
                nn.init.constant_(mod.bias, val=0)

    def forward(self, t, y):
        new_t = t.repeat(y.shape[0], 1)
        yt = torch.cat((y, new_t), 1)
        return self.net(yt - 0.5)


def rk4_alt_step(func, t0, dt, t1, y0):
    """Increment of one 3/8-rule step (torchdiffeq's rk4_alt_step_func)."""
    k1 = func(t0, y0)
    k2 = func(t0 + dt * _one_third, y0 + dt * k1 * _one_third)
    k3 = func(t0 + dt * _two_thirds, y0 + dt * (k2 - k1 * _one_third))
    k4 = func(t1, y0 + dt * (k1 - k2 + k3))
    return (k1 + 3 * (k2 + k3) + k4) * dt * 0.125


def odeint(func, y0, t):
    """Solutions at every point of the 1-D grid ``t`` (stacked on dim 0)."""
    if t.dim() != 1 or t.numel() < 2:
        raise ValueError("odeint: t must be a 1-D grid of at least two times")
    if t.device != y0.device:      # torchdiffeq coerces t to y0's device
        t = t.to(y0.device)
    if bool(t[0] > t[1]):
        t = -t
        base = func

        def func(tt, y):          # noqa: F811
            return -base(-tt, y)
    sol = [y0]
    y = y0
    for i in range(t.numel() - 1):
        t0, t1 = t[i], t[i + 1]
        y = y + rk4_alt_step(func, t0, t1 - t0, t1, y)
        sol.append(y)
    return torch.stack(sol)


class NeuralODE:
    """neuralode_fast.py:35-64."""

    def __init__(self, device=torch.device("cpu")):
        self.timing = torch.from_numpy(np.array([0, 1]).astype("float32")).to(device)
        self.timing_inv = torch.from_numpy(np.array([1, 0]).astype("float32")).to(device)
        self.func = ODEFunc().to(device)
        self.device = device

    def to_device(self, device):
        self.func = self.func.to(device)
        self.timing = self.timing.to(device)
        self.timing_inv = self.timing_inv.to(device)
        self.device = device

    def parameters(self):
        return self.func.parameters()

    def forward(self, u):
        return odeint(self.func, u, self.timing)[1]

    def inverse(self, u):
        return odeint(self.func, u, self.timing_inv)[1]

    def integrate(self, u, t1, t2, device):
        new_time = torch.from_numpy(np.array([t1, t2]).astype("float32")).to(device)
        return odeint(self.func, u, new_time)[1]
