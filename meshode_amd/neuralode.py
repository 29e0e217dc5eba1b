"""NeuralODE deformation (the reference's neuralode_fast.py:7-64 and
cad_neural_deform2.py:40-108, in the configuration of the config-3 workload).

* ``VelocityMLP``: autonomous velocity field v(y) = MLP(y - 0.5),
  3 -> 256 -> 256 -> 256 -> 3 with softplus, Xavier-uniform weights from a
  seed, zero biases.  The reference centres its input the same way
  (neuralode_fast.py:28-32; it also feeds t, which this field does not use).
  The 256x256 layers are FP32 GEMMs (hipBLASLt, on the MFMA units).
* ``odeint_rk4``: fixed-grid RK4 over t in [0, 1] with ``steps`` equal steps
  using the 3/8-rule step that torchdiffeq's ``method="rk4"`` applies
  (the reference calls ``odeint(..., method="rk4")``, neuralode_fast.py:55-60):
      k1 = f(y); k2 = f(y + h k1/3); k3 = f(y + h (k2 - k1/3));
      k4 = f(y + h (k1 - k2 + k3)); y' = y + h (k1 + 3 k2 + 3 k3 + k4) / 8.
* On the GPU, with a 256-wide field in FP32, each forward step of
  ``odeint_rk4`` is one fused HIP kernel (``mo_flow256_step``: the four
  stages, the 256 x 256 layers on FP32 MFMA, softplus in their epilogue);
  the parameters are packed once per call (``mo_flow256_pack``).  The
  adjoint's recomputation saves each step's activations
  (``mo_flow256_step_save``), a second kernel runs the backward through the
  four stages (``mo_flow256_step_backward``, the transposed 256 x 256
  products on MFMA), and the parameter gradients are GEMMs over the saved
  tensors.
* ``RK4Adjoint``: the gradient of the discretized flow (discrete adjoint).
  The forward keeps only the state at each step boundary (one checkpoint
  per step); the backward recomputes one step at a time from its
  checkpoint and back-propagates through its four stages, so memory is one
  step's activations instead of all ``4 * steps`` evaluations.
* ``DistanceRigidityLoss``: 1/2 sum d(y)^2 over all pushed points (the
  pyDeform distance kernels) + 1/2 rho^2 sum ||(y_b - y_a) - (y0_b - y0_a)||^2
  over the source mesh edges (the graph-edge kernels).
"""
from __future__ import annotations

import math

import torch
from torch import nn
from torch.autograd import Function

from . import pyDeform


class VelocityMLP(nn.Module):
    def __init__(self, hidden=256, seed=0, dtype=torch.float32):
        super().__init__()
        g = torch.Generator().manual_seed(seed)
        dims = [3, hidden, hidden, hidden, 3]
        self.weights = nn.ParameterList()
        self.biases = nn.ParameterList()
        for a, b in zip(dims[:-1], dims[1:]):
            bound = math.sqrt(6.0 / (a + b))
            self.weights.append(nn.Parameter((torch.rand(b, a, generator=g, dtype=torch.float64) * 2 - 1)
                                             .mul_(bound).to(dtype)))
            self.biases.append(nn.Parameter(torch.zeros(b, dtype=dtype)))

    def forward(self, y):
        h = y - 0.5
        n = len(self.weights)
        for i, (W, b) in enumerate(zip(self.weights, self.biases)):
            h = torch.addmm(b, h, W.t())
            if i < n - 1:
                h = nn.functional.softplus(h)
        return h


def rk4_step(f, y, h):
    """One 3/8-rule RK4 step (torchdiffeq's ``rk4`` fixed-grid step)."""
    k1 = f(y)
    k2 = f(y + h * k1 / 3.0)
    k3 = f(y + h * (k2 - k1 / 3.0))
    k4 = f(y + h * (k1 - k2 + k3))
    return y + h * (k1 + 3.0 * (k2 + k3) + k4) * 0.125


def _fused_field(func, y):
    """True when the fused MFMA step applies: a 256-wide FP32 VelocityMLP on
    the GPU."""
    return (y.is_cuda and y.dtype == torch.float32 and isinstance(func, VelocityMLP) and len(func.weights) == 4
            and tuple(func.weights[1].shape) == (256, 256) and func.weights[0].shape[1] == 3)


def flow256_pack(func):
    """The field's parameters in the fused step's packed layout (device)."""
    import ctypes as C
    from . import _lib
    ts = [t.detach().contiguous() for pair in zip(func.weights, func.biases) for t in pair]
    out = torch.empty(int(_lib.load().mo_flow256_packed_floats()), dtype=torch.float32, device=ts[0].device)
    _lib.call("mo_flow256_pack", *[C.c_void_p(t.data_ptr()) for t in ts], C.c_void_p(out.data_ptr()),
              pyDeform._stream())
    return out


def flow256_step(packed, y, h):
    """One fused RK4 step (FP32, GPU) of the packed 256-wide field."""
    import ctypes as C
    from . import _lib
    y = y.contiguous()
    y1 = torch.empty_like(y)
    _lib.call("mo_flow256_step", C.c_void_p(y.data_ptr()), y.shape[0], C.c_void_p(packed.data_ptr()), float(h),
              C.c_void_p(y1.data_ptr()), pyDeform._stream())
    return y1


class RK4Adjoint(Function):
    """y1 = Phi(y0; theta) for ``steps`` RK4 steps; backward by recomputation
    from per-step checkpoints."""

    @staticmethod
    def forward(ctx, y0, steps, func, *params):
        h = 1.0 / steps
        ys = [y0.detach()]
        with torch.no_grad():
            y = y0.detach()
            packed = flow256_pack(func) if _fused_field(func, y) else None
            for _ in range(steps):
                y = flow256_step(packed, y, h) if packed is not None else rk4_step(func, y, h)
                ys.append(y)
        ctx.func, ctx.steps, ctx.h, ctx.packed = func, steps, h, packed
        ctx.save_for_backward(*ys[:-1])
        ctx.n_params = len(params)
        return y

    @staticmethod
    def _fused_backward(ctx, gy):
        """Per step, last first: the fused step again, saving its stage
        inputs and activations (mo_flow256_step_save), the fused backward
        through its four stages (mo_flow256_step_backward: dL/dy and the
        pre-activation gradients), and the parameter gradients as sums over
        the step's 4 n stage points: a split-K MFMA kernel for W1 and W2
        (mo_flow256_dw_grads), one reduction kernel for W0, W3 and the
        biases (mo_flow256_small_grads)."""
        import ctypes as C
        from . import _lib
        ys = ctx.saved_tensors
        n, dev, h = ys[0].shape[0], ys[0].device, ctx.h
        params = list(ctx.func.parameters())
        gp = [torch.zeros_like(p) for p in params]
        Hs = torch.empty(3 * 4 * n * 256, dtype=torch.float32, device=dev)
        dAs = torch.empty_like(Hs)
        Xs = torch.empty(4 * n * 3, dtype=torch.float32, device=dev)
        Gs = torch.empty_like(Xs)
        y1 = torch.empty(n, 3, dtype=torch.float32, device=dev)
        nsmall = int(_lib.load().mo_flow256_small_grad_floats())
        small = torch.empty(nsmall, dtype=torch.float32, device=dev)
        work = torch.empty(256 * nsmall, dtype=torch.float32, device=dev)
        dw = torch.empty(2, 256, 256, dtype=torch.float32, device=dev)
        dwork = torch.empty(int(_lib.load().mo_flow256_dw_work_floats()), dtype=torch.float32, device=dev)
        H, dA = Hs.view(3, 4 * n, 256), dAs.view(3, 4 * n, 256)
        X, G = Xs.view(4 * n, 3), Gs.view(4 * n, 3)
        g = gy.contiguous().to(torch.float32)
        ptr = lambda t: C.c_void_p(t.data_ptr())      # noqa: E731
        for k in reversed(range(ctx.steps)):
            _lib.call("mo_flow256_step_save", ptr(ys[k].contiguous()), n, ptr(ctx.packed), float(h), ptr(y1),
                      ptr(Hs), ptr(Xs), pyDeform._stream())
            gnew = torch.empty(n, 3, dtype=torch.float32, device=dev)
            _lib.call("mo_flow256_step_backward", n, ptr(ctx.packed), float(h), ptr(Hs), ptr(g), ptr(gnew),
                      ptr(dAs), ptr(Gs), pyDeform._stream())
            # weights[0..3], biases[0..3]: W0 [256,3], W1, W2 [256,256], W3 [3,256]
            _lib.call("mo_flow256_dw_grads", n, ptr(dAs), ptr(Hs), ptr(dw), ptr(dwork), pyDeform._stream())
            gp[1].add_(dw[0])
            gp[2].add_(dw[1])
            _lib.call("mo_flow256_small_grads", n, ptr(dAs), ptr(Hs), ptr(Xs), ptr(Gs), ptr(small), ptr(work),
                      pyDeform._stream())
            gp[0].add_(small[0:768].view(256, 3))
            gp[4].add_(small[768:1024])
            gp[5].add_(small[1024:1280])
            gp[6].add_(small[1280:1536])
            gp[3].add_(small[1536:2304].view(3, 256))
            gp[7].add_(small[2304:2307])
            g = gnew
        return (g, None, None, *gp)

    @staticmethod
    def backward(ctx, gy):
        if ctx.packed is not None:
            return RK4Adjoint._fused_backward(ctx, gy)
        ys = ctx.saved_tensors
        params = [p for p in ctx.func.parameters()]
        gp = [torch.zeros_like(p) for p in params]
        for n in reversed(range(ctx.steps)):
            with torch.enable_grad():
                y = ys[n].detach().requires_grad_(True)
                y1 = rk4_step(ctx.func, y, ctx.h)
                grads = torch.autograd.grad(y1, [y] + params, gy, allow_unused=True)
            gy = grads[0]
            for i, g in enumerate(grads[1:]):
                if g is not None:
                    gp[i] += g
        return (gy, None, None, *gp)


def odeint_rk4(func: nn.Module, y0, steps=16):
    """Integrate dy/dt = func(y) over [0, 1] with ``steps`` RK4 steps."""
    return RK4Adjoint.apply(y0, steps, func, *func.parameters())


class _DistanceLoss(Function):
    @staticmethod
    def forward(ctx, y, pid):
        ctx.save_for_backward(y)
        ctx.pid = pid
        return (pyDeform.DistanceFieldLoss_forward(y.contiguous(), pid) * 0.5).sum()

    @staticmethod
    def backward(ctx, g):
        (y,) = ctx.saved_tensors
        return g * pyDeform.DistanceFieldLoss_backward(y.contiguous(), ctx.pid), None


class _GraphLoss(Function):
    @staticmethod
    def forward(ctx, y, E, pid):
        ctx.save_for_backward(y, E)
        ctx.pid = pid
        return (pyDeform.GraphEdgeLoss_forward(y.contiguous(), E, pid) * 0.5).sum()

    @staticmethod
    def backward(ctx, g):
        y, E = ctx.saved_tensors
        return g * pyDeform.GraphEdgeLoss_backward(y.contiguous(), E, ctx.pid), None, None


class NeuralODEDeformer(nn.Module):
    """Pushes ``points`` (the first ``n_mesh`` are the source mesh vertices,
    with edges ``E``) through the flow; loss = distance to the target grid
    (all points) + rho^2 * rigidity (mesh edges)."""

    def __init__(self, points, E, pid, n_mesh, rho=0.1, steps=16, hidden=256, seed=0):
        super().__init__()
        self.field = VelocityMLP(hidden, seed).to(points.device)
        self.register_buffer("y0", points.detach().contiguous())
        self.E = E.contiguous()
        self.pid = pid
        self.n_mesh = n_mesh
        self.rho2 = rho * rho
        self.steps = steps
        pyDeform.StoreGraphInformation(self.y0[:n_mesh].contiguous(), self.E, pid)

    def forward(self):
        y1 = odeint_rk4(self.field, self.y0, self.steps)
        lossD = _DistanceLoss.apply(y1, self.pid)
        lossR = _GraphLoss.apply(y1[: self.n_mesh], self.E, self.pid)
        return lossD + self.rho2 * lossR, y1
