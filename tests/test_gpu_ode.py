"""GPU parity of the fused NeuralODE step (mo_ode_rk4_forward / _backward):
one 3/8-rule RK4 step of ODEFunc (neuralode_fast.py:7-64) and its gradient.

The kernels compute in FP32 with their own summation order; the bar is the
FP64 autograd of the same step (torch on the CPU, same weights): every
tensor (y1, dL/dy0, each parameter's gradient) within 1e-4 of its largest
magnitude, or within 1.5 times the error of the float32 torch step on the
CPU (the reference's own precision) where that is larger.  A pre-activation within
FP32 rounding of a LeakyReLU kink takes the other slope in either float32
computation: for dL/dy0 at most 1 point in 10^4 may exceed the bar, by at
most 1e-2 of the scale.  (On the interval [0.2, 0.7] at 40000 points one
such kink is met: the fused kernel and the float32 torch step then differ
from FP64 by the same 2.5e-3 in dL/dW2.)
Cases: the forward grid [0, 1], the inverse grid [1, 0] and a general
interval; point counts 1 (a partial chunk), 64 (one chunk), 1000 and 40000
(more chunks than backward blocks: the block-stride path); y0 inside and
outside [0, 1]^3.  The backward is deterministic (repeat calls are
bit-identical); NeuralODE.forward / inverse on the GPU take the fused path
and match the CPU module.
"""
import numpy as np
import pytest
import torch

from meshode_amd import _lib
from meshode_amd.layers import neuralode_fast as nf

pytestmark = pytest.mark.gpu

TOL = 1e-4


def _say(msg):
    print(msg, flush=True)


def _reference(func, y0, grid, w):
    """FP64 CPU autograd of the torch step: y1, dL/dy0, dL/dparams for L = sum(w * y1)."""
    f64 = nf.ODEFunc().double()
    f64.load_state_dict({k: v.double().cpu() for k, v in func.state_dict().items()})
    y = y0.detach().double().cpu().requires_grad_(True)
    y1 = nf.odeint(f64, y, torch.tensor(grid, dtype=torch.float64))[1]
    (y1 * w.double().cpu()).sum().backward()
    return y1.detach(), y.grad, torch.cat([p.grad.reshape(-1) for p in f64.parameters()])


def _close(got, want, what, ref_err, per_point=False):
    """got within max(TOL * scale, 1.5 ref_err) of want, ref_err being the
    float32 torch step's own error on the same tensor."""
    got = got.detach().double().cpu()
    scale = max(float(want.abs().max()), 1e-30)
    diff = (got - want).abs()
    err = float(diff.max())
    bar = max(TOL * scale, 1.5 * ref_err)
    _say(f"[ode]   {what}: max |diff| {err:.3e} (scale {scale:.3e}; torch f32 {ref_err:.3e})")
    if not per_point:
        assert err <= bar, what
        return
    # a point whose pre-activation lies within FP32 rounding of a LeakyReLU
    # kink takes the other slope (as the float32 torch step may): at most 1
    # point in 10^4 may exceed the bar, by at most 1e-2 of the scale
    bad = int((diff.reshape(diff.shape[0], -1).max(1).values > bar).sum())
    assert bad <= max(1, diff.shape[0] // 10000) and err <= max(1e-2 * scale, ref_err), (what, bad)


@pytest.mark.parametrize("grid", [(0.0, 1.0), (1.0, 0.0), (0.2, 0.7)])
@pytest.mark.parametrize("n", [1, 64, 1000, 40000])
def test_fused_step_against_fp64(gpu, grid, n):
    _say(f"[ode] fused RK4 step grid={grid} n={n}")
    torch.manual_seed(n + int(10 * grid[0]))
    func = nf.ODEFunc().to(gpu)
    y0 = (torch.rand(n, 3) * 1.4 - 0.2).to(gpu).requires_grad_(True)
    w = torch.randn(n, 3).to(gpu)
    y1 = nf.odeint(func, y0, torch.tensor(grid, dtype=torch.float32, device=gpu))[1]
    (y1 * w).sum().backward()
    ry1, rgy0, rgp = _reference(func, y0, grid, w)
    # the float32 torch step on the CPU, same weights: the reference's own precision
    f32 = nf.ODEFunc()
    f32.load_state_dict({k: v.cpu() for k, v in func.state_dict().items()})
    yc = y0.detach().cpu().requires_grad_(True)
    cy1 = nf.odeint(f32, yc, torch.tensor(grid))[1]
    (cy1 * w.cpu()).sum().backward()
    cgp = torch.cat([p.grad.reshape(-1) for p in f32.parameters()]).double()

    def ref_err(a, b):
        return float((a.detach().double() - b).abs().max())

    _close(y1, ry1, "y1", ref_err(cy1, ry1))
    _close(y0.grad, rgy0, "dL/dy0", ref_err(yc.grad, rgy0), per_point=True)
    gp = torch.cat([p.grad.reshape(-1) for p in func.parameters()])
    names = ["W0", "b0", "W1", "b1", "W2", "b2", "W3", "b3"]
    off = 0
    for name, p in zip(names, func.parameters()):
        k = p.numel()
        _close(gp[off:off + k], rgp[off:off + k], "dL/d" + name, ref_err(cgp[off:off + k], rgp[off:off + k]))
        off += k


def test_fused_backward_is_deterministic_and_y0_optional(gpu):
    _say("[ode] repeat backward bit-identical; gy0 omitted")
    torch.manual_seed(5)
    func = nf.ODEFunc().to(gpu)
    y0 = torch.rand(5000, 3, device=gpu)
    w = torch.randn(5000, 3, device=gpu)
    grads = []
    for _ in range(2):
        func.zero_grad()
        (nf.fused_step(func, y0, 1.0, 0.0) * w).sum().backward()      # y0 needs no gradient
        grads.append(torch.cat([p.grad.reshape(-1) for p in func.parameters()]))
    assert torch.equal(grads[0], grads[1])
    yr = y0.clone().requires_grad_(True)
    func.zero_grad()
    (nf.fused_step(func, yr, 1.0, 0.0) * w).sum().backward()
    assert torch.equal(torch.cat([p.grad.reshape(-1) for p in func.parameters()]), grads[0])


def test_neuralode_module_uses_fused_step(gpu, monkeypatch):
    _say("[ode] NeuralODE forward/inverse on the GPU against the CPU module")
    torch.manual_seed(7)
    m = nf.NeuralODE(gpu)
    cpu = nf.NeuralODE(torch.device("cpu"))
    cpu.func.load_state_dict({k: v.cpu() for k, v in m.func.state_dict().items()})
    u = torch.rand(3000, 3)
    calls = []
    orig = nf._FusedRK4Step.apply
    monkeypatch.setattr(nf._FusedRK4Step, "apply", lambda *a: calls.append(1) or orig(*a))
    a = m.forward(u.to(gpu))
    b = m.inverse(u.to(gpu))
    c = m.integrate(u.to(gpu), 0.3, 0.6, gpu)
    monkeypatch.undo()
    assert len(calls) == 3
    for got, want in ((a, cpu.forward(u)), (b, cpu.inverse(u)), (c, cpu.integrate(u, 0.3, 0.6, "cpu"))):
        assert float((got.cpu() - want).abs().max()) <= 1e-5
    with torch.no_grad():
        assert torch.equal(m.forward(u.to(gpu)), a.detach())


def test_fused_step_errors_and_empty(gpu):
    _say("[ode] empty input and argument errors")
    func = nf.ODEFunc().to(gpu)
    e = torch.empty(0, 3, device=gpu, requires_grad=True)
    y = nf.fused_step(func, e, 0.0, 1.0)
    y.sum().backward()
    assert y.shape == (0, 3)
    assert all(float(p.grad.abs().max()) == 0.0 for p in func.parameters())
    with pytest.raises(TypeError):
        nf.fused_step(func, torch.zeros(4, 3, dtype=torch.float64, device=gpu), 0.0, 1.0)
    import ctypes as C
    with pytest.raises(_lib.MoError, match="bad arguments"):
        _lib.call("mo_ode_rk4_forward", None, 5, None, (C.c_float * 4)(), 1.0, 1.0, None, None, None)
