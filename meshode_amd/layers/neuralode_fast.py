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
  Gradients are those of plain autograd through the step, as with
  torchdiffeq's non-adjoint ``odeint``.

On a GPU, ``odeint`` of an ``ODEFunc`` runs each grid interval as one fused
HIP step (``mo_ode_rk4_forward`` / ``mo_ode_rk4_backward`` in libmeshode:
the four stages and the MLP in one kernel, the backward through all four
stages and the parameter gradient in another) behind a
``torch.autograd.Function``; the torch expression above is the CPU path.
"""
from __future__ import annotations

import ctypes as C
import functools

import numpy as np
import torch
from torch import nn

from .. import _lib

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


N_PARAMS = 5503     # 4*50 + 50 + 50*50 + 50 + 50*50 + 50 + 3*50 + 3


@functools.lru_cache(maxsize=64)
def step_times(t0, t1):
    """(tau[4], sign, dt) of one 3/8-rule step from t0 to t1, in float32 as
    the torch step forms them: a decreasing grid runs as t -> -t with
    f -> -f(-t, y), so the network sees tau = sign * stage time."""
    tt = torch.tensor([float(t0), float(t1)], dtype=torch.float32)
    sign = 1.0
    if bool(tt[0] > tt[1]):
        tt = -tt
        sign = -1.0
    a, b = tt[0], tt[1]
    dt = b - a
    stages = [a, a + dt * _one_third, a + dt * _two_thirds, b]
    tau = tuple(float(x) if sign > 0 else float(-x) for x in stages)
    return tau, sign, float(dt)


class _FusedRK4Step(torch.autograd.Function):
    """y1 = one RK4 step of ODEFunc from y0 (HIP kernels); flat = the
    module's parameters concatenated in parameter order."""

    @staticmethod
    def forward(ctx, y0, flat, tau, sign, dt, need_grad):
        from .. import pyDeform
        y0c = y0.detach().contiguous()
        n = y0c.shape[0]
        y1 = torch.empty_like(y0c)
        zsave = torch.empty((n, 9), dtype=torch.float32, device=y0c.device) if need_grad else None
        taus = (C.c_float * 4)(*tau)
        _lib.call("mo_ode_rk4_forward", pyDeform._ptr(y0c), n, C.c_void_p(flat.data_ptr()), taus, sign, dt,
                  pyDeform._ptr(y1), pyDeform._ptr(zsave), pyDeform._stream())
        if need_grad:
            ctx.save_for_backward(y0c, flat, zsave)
        ctx.tau, ctx.sign, ctx.dt = tau, sign, dt
        return y1

    @staticmethod
    def backward(ctx, gy1):
        from .. import pyDeform
        y0c, flat, zsave = ctx.saved_tensors
        n = y0c.shape[0]
        gy1 = gy1.contiguous()
        gy0 = torch.empty_like(y0c) if ctx.needs_input_grad[0] else None
        gflat = torch.empty(N_PARAMS, dtype=torch.float32, device=y0c.device)
        nw = int(_lib.load().mo_ode_rk4_workspace_bytes(n))
        work = torch.empty(max(nw // 8, 1), dtype=torch.float64, device=y0c.device)
        taus = (C.c_float * 4)(*ctx.tau)
        _lib.call("mo_ode_rk4_backward", pyDeform._ptr(y0c), n, C.c_void_p(flat.data_ptr()), taus, ctx.sign, ctx.dt,
                  pyDeform._ptr(zsave), pyDeform._ptr(gy1), pyDeform._ptr(gy0), C.c_void_p(gflat.data_ptr()),
                  pyDeform._ptr(work), nw, pyDeform._stream())
        return gy0, gflat, None, None, None, None


def fused_step(func, y0, t0, t1):
    """One RK4 step of the ODEFunc ``func`` from t0 to t1 on the GPU."""
    if y0.dtype != torch.float32 or y0.dim() != 2 or y0.shape[1] != 3:
        raise TypeError("fused_step: y0 must be float32 [n,3]")
    params = list(func.parameters())
    flat = torch.cat([p.reshape(-1) for p in params])
    if flat.numel() != N_PARAMS:
        raise ValueError("fused_step: func is not the 4-50-50-50-3 ODEFunc")
    tau, sign, dt = step_times(t0, t1)
    need = torch.is_grad_enabled() and (y0.requires_grad or flat.requires_grad)
    return _FusedRK4Step.apply(y0, flat, tau, sign, dt, need)


def odeint(func, y0, t):
    """Solutions at every point of the 1-D grid ``t`` (stacked on dim 0)."""
    if t.dim() != 1 or t.numel() < 2:
        raise ValueError("odeint: t must be a 1-D grid of at least two times")
    if t.device != y0.device:      # torchdiffeq coerces t to y0's device
        t = t.to(y0.device)
    if y0.is_cuda and isinstance(func, ODEFunc):
        tl = t.tolist()
        sol = [y0]
        y = y0
        for i in range(len(tl) - 1):
            y = fused_step(func, y, tl[i], tl[i + 1])
            sol.append(y)
        return torch.stack(sol)
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
        if u.is_cuda:          # the grid [0, 1] without reading it back from the device
            return fused_step(self.func, u, 0.0, 1.0)
        return odeint(self.func, u, self.timing)[1]

    def inverse(self, u):
        if u.is_cuda:
            return fused_step(self.func, u, 1.0, 0.0)
        return odeint(self.func, u, self.timing_inv)[1]

    def integrate(self, u, t1, t2, device):
        new_time = torch.from_numpy(np.array([t1, t2]).astype("float32")).to(device)
        return odeint(self.func, u, new_time)[1]
