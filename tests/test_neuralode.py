"""NeuralODE flow (config 3): the 3/8-rule RK4 integrator, the checkpointed
discrete adjoint and the full deformer loss.

CPU (FP64, pure torch): one step on a linear field equals the degree-4 Taylor
polynomial of exp(hA) (exact for 4-stage 4th-order RK); one step on a
nonlinear field equals a generic explicit-RK evaluation of the 3/8-rule
Butcher tableau; the adjoint equals autograd through the unrolled steps and
central differences.  GPU: FP32 forward and gradients against FP64, and the
deformer's loss against the oracle's distance and edge losses."""
import numpy as np
import pytest
import torch

from meshode_amd import neuralode as NO


def test_linear_field_taylor():
    torch.manual_seed(0)
    A = torch.randn(3, 3, dtype=torch.float64) * 0.7
    f = lambda y: y @ A.t()        # noqa: E731
    y = torch.randn(5, 3, dtype=torch.float64)
    h = 0.3
    M = h * A
    P = torch.eye(3, dtype=torch.float64) + M + M @ M / 2 + M @ M @ M / 6 + M @ M @ M @ M / 24
    assert torch.allclose(NO.rk4_step(f, y, h), y @ P.t(), rtol=1e-14, atol=1e-14)


def test_three_eighths_tableau():
    """Generic explicit RK with Kutta's 3/8-rule tableau."""
    c = [0.0, 1 / 3, 2 / 3, 1.0]
    a = [[], [1 / 3], [-1 / 3, 1.0], [1.0, -1.0, 1.0]]
    b = [1 / 8, 3 / 8, 3 / 8, 1 / 8]
    f = lambda y: torch.sin(3 * y) + y * y      # noqa: E731
    y = torch.tensor([[0.3, -0.7, 1.1]], dtype=torch.float64)
    h = 0.4
    k = []
    for i in range(4):
        yi = y + h * sum(a[i][j] * k[j] for j in range(i)) if i else y
        k.append(f(yi))
    expect = y + h * sum(bi * ki for bi, ki in zip(b, k))
    assert torch.allclose(NO.rk4_step(f, y, h), expect, rtol=1e-15, atol=1e-15)


def _tiny(seed=1):
    m = NO.VelocityMLP(hidden=8, seed=seed, dtype=torch.float64)
    y0 = torch.rand(7, 3, dtype=torch.float64, generator=torch.Generator().manual_seed(seed))
    return m, y0


def test_adjoint_equals_unrolled_autograd():
    m, y0 = _tiny()
    y0a = y0.clone().requires_grad_(True)
    out = NO.odeint_rk4(m, y0a, steps=4)
    w = torch.linspace(-1, 1, out.numel(), dtype=torch.float64).view_as(out)
    (out * w).sum().backward()
    ga = [y0a.grad] + [p.grad.clone() for p in m.parameters()]
    for p in m.parameters():
        p.grad = None
    y0b = y0.clone().requires_grad_(True)
    y = y0b
    for _ in range(4):
        y = NO.rk4_step(m, y, 0.25)
    (y * w).sum().backward()
    gb = [y0b.grad] + [p.grad for p in m.parameters()]
    assert torch.allclose(out, y.detach(), rtol=0, atol=0)
    for a_, b_ in zip(ga, gb):
        assert torch.allclose(a_, b_, rtol=1e-12, atol=1e-14)


def test_adjoint_central_differences():
    m, y0 = _tiny(2)
    w = torch.randn(7, 3, dtype=torch.float64, generator=torch.Generator().manual_seed(3))

    def loss():
        return (NO.odeint_rk4(m, y0, steps=3) * w).sum()

    l0 = loss()
    l0.backward()
    W1 = m.weights[1]
    g = W1.grad.clone()
    eps = 1e-6
    for (i, j) in [(0, 0), (3, 5), (7, 2)]:
        with torch.no_grad():
            W1[i, j] += eps
        lp = loss().item()
        with torch.no_grad():
            W1[i, j] -= 2 * eps
        lm = loss().item()
        with torch.no_grad():
            W1[i, j] += eps
        assert g[i, j].item() == pytest.approx((lp - lm) / (2 * eps), rel=1e-6, abs=1e-10)


@pytest.mark.gpu
def test_gpu_flow_and_loss_against_fp64(gpu, oracle):
    from meshode_amd import pyDeform, synth
    print("[ode] case=box1000_points2000", flush=True)
    sV, sF, tV, tF = synth.box_union_pair(71, n_boxes=4, n_verts=1000)
    N = 32
    pid = pyDeform.InitializeDeformTemplate(torch.from_numpy(tV.astype(np.float32)), torch.from_numpy(tF), 0, N)
    grid = pyDeform.TemplateGrid(pid)[0].cpu().numpy().reshape(-1)
    V = torch.from_numpy(sV.astype(np.float32))
    pyDeform.NormalizeByTemplate(V, pid)
    # points: mesh vertices + random points near the surface
    rng = np.random.default_rng(0)
    extra = (V.numpy()[rng.integers(0, len(V), 1000)] + rng.normal(0, 0.01, (1000, 3))).astype(np.float32)
    pts = torch.cat([V, torch.from_numpy(extra)]).cuda()
    E = torch.from_numpy(np.concatenate([sF[:, [0, 1]], sF[:, [1, 2]], sF[:, [2, 0]]]).astype(np.int32)).cuda()
    model = NO.NeuralODEDeformer(pts, E, pid, n_mesh=len(V), rho=0.1, steps=16, seed=5)
    loss, y1 = model()
    loss.backward()
    # FP64 reference of the flow with the same parameters
    m64 = NO.VelocityMLP(hidden=256, seed=5, dtype=torch.float64)
    with torch.no_grad():
        y64 = NO.odeint_rk4(m64, pts.double().cpu(), steps=16)
    assert torch.allclose(y1.detach().double().cpu(), y64, rtol=0, atol=2e-5)
    # loss against the oracle's layers evaluated on the GPU result
    y = y1.detach().cpu().numpy()
    d = oracle.distance_loss_fwd(grid, N, y)
    off = oracle.store_graph(V.numpy(), E.cpu().numpy())
    r = oracle.graph_fwd(y[: len(V)], E.cpu().numpy(), off)
    ref = 0.5 * d.astype(np.float64).sum() + 0.01 * 0.5 * r.astype(np.float64).sum()
    assert loss.item() == pytest.approx(ref, rel=1e-4)
    # parameter gradients against the FP64 adjoint with the oracle's loss gradient
    gd = oracle.distance_loss_bwd(grid, N, y).astype(np.float64)
    gr = oracle.graph_bwd(y[: len(V)], E.cpu().numpy(), off).astype(np.float64)
    gy = gd.copy()
    gy[: len(V)] += 0.01 * gr
    y0 = pts.double().cpu()
    out = NO.odeint_rk4(m64, y0, steps=16)
    out.backward(torch.from_numpy(gy))
    for p32, p64 in zip(model.field.parameters(), m64.parameters()):
        a, b = p32.grad.double().cpu(), p64.grad
        assert torch.linalg.norm(a - b) <= 2e-3 * torch.linalg.norm(b) + 1e-9


@pytest.mark.gpu
@pytest.mark.parametrize("n", [1, 64, 100, 55000])
def test_gpu_fused_flow256_step_against_fp64(gpu, n):
    """mo_flow256_step (the four RK4 stages of the 256-wide field on FP32
    MFMA) against the FP64 torch step with the same weights: y1 within 1e-5
    of the FP64 step's increment scale, or within 1.5 times the float32 torch
    step's own error where that is larger; points inside and outside the
    unit box."""
    print(f"[ode] fused flow256 step n={n}", flush=True)
    torch.manual_seed(n)
    m = NO.VelocityMLP(hidden=256, seed=3).to(gpu)
    with torch.no_grad():                      # non-zero biases exercise the epilogue
        for b in m.biases:
            b.normal_(0, 0.3)
    y = (torch.rand(n, 3) * 1.6 - 0.3).to(gpu)
    h = 1.0 / 16
    got = NO.flow256_step(NO.flow256_pack(m), y, h)
    m64 = NO.VelocityMLP(hidden=256, seed=3, dtype=torch.float64)
    m64.load_state_dict({k: v.double().cpu() for k, v in m.state_dict().items()})
    with torch.no_grad():
        want = NO.rk4_step(m64, y.double().cpu(), h)
        f32 = NO.rk4_step(m.cpu(), y.cpu(), h).double()
    m.to(gpu)
    scale = float((want - y.double().cpu()).abs().max())
    err = float((got.double().cpu() - want).abs().max())
    ref = float((f32 - want).abs().max())
    print(f"[ode]   max |diff| {err:.3e} (increment scale {scale:.3e}; torch f32 {ref:.3e})", flush=True)
    assert err <= max(1e-5 * scale, 1.5 * ref)


@pytest.mark.gpu
def test_gpu_odeint_uses_fused_step(gpu, monkeypatch):
    """odeint_rk4's forward takes the fused step for a 256-wide FP32 field on
    the GPU (one kernel per step) and the torch step otherwise."""
    m = NO.VelocityMLP(hidden=256, seed=1).to(gpu)
    calls = []
    orig = NO.flow256_step
    monkeypatch.setattr(NO, "flow256_step", lambda *a: calls.append(1) or orig(*a))
    y = torch.rand(300, 3, device=gpu)
    out = NO.odeint_rk4(m, y, steps=4)
    assert len(calls) == 4 and out.shape == y.shape
    small = NO.VelocityMLP(hidden=8, seed=1).to(gpu)
    NO.odeint_rk4(small, y, steps=4)
    assert len(calls) == 4


@pytest.mark.gpu
@pytest.mark.parametrize("n", [40, 1000])
def test_gpu_fused_adjoint_against_fp64(gpu, n):
    """The fused adjoint (mo_flow256_step_save / _backward + GEMMs) of a
    4-step integration against FP64 autograd through the same steps: the
    parameter gradients and dL/dy0 within 1e-4 (relative, in norm); 40
    points exercise the one-block (at most 32 points) path."""
    print(f"[ode] fused adjoint n={n}", flush=True)
    torch.manual_seed(n)
    m = NO.VelocityMLP(hidden=256, seed=7).to(gpu)
    with torch.no_grad():
        for b in m.biases:
            b.normal_(0, 0.2)
    y0 = (torch.rand(n, 3) * 1.2 - 0.1).to(gpu).requires_grad_(True)
    w = torch.randn(n, 3).to(gpu)
    out = NO.odeint_rk4(m, y0, steps=4)
    (out * w).sum().backward()
    m64 = NO.VelocityMLP(hidden=256, seed=7, dtype=torch.float64)
    m64.load_state_dict({k: v.double().cpu() for k, v in m.state_dict().items()})
    y64 = y0.detach().double().cpu().requires_grad_(True)
    y = y64
    for _ in range(4):
        y = NO.rk4_step(m64, y, 0.25)
    (y * w.double().cpu()).sum().backward()
    assert float((out.detach().double().cpu() - y.detach()).abs().max()) <= 1e-5 * float(y.detach().abs().max())
    pairs = [(y0.grad, y64.grad, "dL/dy0")] + [(a.grad, b.grad, name) for (name, a), b in
                                            zip(m.named_parameters(), m64.parameters())]
    for got, want, name in pairs:
        rel = float(torch.linalg.norm(got.double().cpu() - want) / torch.linalg.norm(want))
        print(f"[ode]   {name}: relative error {rel:.2e}", flush=True)
        assert rel <= 1e-4, name
