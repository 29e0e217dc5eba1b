"""GPU parity at the benchmark configurations' own workloads.

* headline (5000-vertex box unions vs their jittered targets, 64^3): 16
  seeds, each GPU result's energy (the oracle's FP64 cost, Deformer::Deform,
  deformer.cc:18-92) within 1e-3 of the oracle LM's final energy;
* config 1 (icosphere L4 source vs displaced L5 target scaled (1, 0.7, 1.3),
  64^3, lambda 1, rigid_deform.cc:49-61): 4 seeds, same bar, on the default
  register/LDS kernel;
* config 2 (RigidLossLayer on the 5000-vertex box union at 128^3,
  rigid_loss_layer.py:7-44): layer outputs and gradient bit-exact against the
  oracle's layers on the GPU's grid; the gradient against FP64 central
  differences (h = 1e-4) on 64 vertices away from cell faces and the 0.2
  truncation; 20 Adam steps (rigid_deform.py:25-44) with the gradient
  bit-exact at every iterate, and the same loop on the CPU driven by the
  oracle's layers;
* config 4 (icosphere levels 3-6, 64^3 / 128^3, lambda in [0.5, 2]): the
  level-6 pair (40962 vertices, 128^3) on the register/LDS team kernel against
  the oracle's converged energy (golden fixture), and a mixed batch whose
  every pair is bit-identical alone and in the mix."""
import os
from concurrent.futures import ThreadPoolExecutor

import numpy as np
import pytest
import torch

from meshode_amd import deform, pyDeform, synth

pytestmark = pytest.mark.gpu

ENERGY_RTOL = 1e-3
GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def _say(msg):
    print(msg, flush=True)


def _oracle_energies(oracle, pairs, N, lams, outs):
    """For each pair: the oracle's converged LM energy on the GPU template's
    FP64 grid (bit-exact to the oracle's own grid) and the GPU result's
    energy under the oracle's FP64 cost."""
    prep = []
    for (sV, sF, tV, tF), lam, out in zip(pairs, lams, outs):
        pid = deform.template_f64(tV, tF, N)
        _, scale, pos = pyDeform.TemplateInfo(pid)
        grid = pyDeform.TemplateGrid(pid)[0].cpu().numpy().reshape(-1)
        _destroy(pid)
        sn = oracle.apply_transform(sV, scale, pos)
        prep.append((grid, sn, oracle.face_edges(sF), lam, (out.cpu().numpy() - pos) / scale))

    def one(p):
        grid, sn, E, lam, xg = p
        _, so, _ = oracle.lm_solve(grid, N, sn, E, lam)
        eg, _, _ = oracle.lm_cost(grid, N, xg, sn, E, lam)
        return so, eg

    with ThreadPoolExecutor(max_workers=min(16, len(prep))) as ex:
        return list(ex.map(one, prep))


def _destroy(pid):
    from meshode_amd import _lib
    _lib.call("mo_template_destroy", int(pid))


def test_headline_sixteen_seeds(gpu, oracle):
    _say("[cfg] headline: 16 box-union seeds, 5000 verts, 64^3")
    pairs = [synth.box_union_pair(s) for s in range(16)]
    outs, st = deform.rigid_deform_batch(pairs, grid_resolution=64, lambdas=1.0)
    assert all(s["solver_path_name"] == "lds" for s in st)
    res = _oracle_energies(oracle, pairs, 64, [1.0] * 16, outs)
    worst = 0.0
    for s, ((so, eg), gs) in enumerate(zip(res, st)):
        rel = (eg - so["final_cost"]) / so["final_cost"]
        worst = max(worst, abs(rel))
        _say(f"[cfg] seed {s}: oracle E={so['final_cost']:.8g} it={so['iterations']}  gpu E={eg:.8g} "
             f"it={gs['iterations']}  rel {rel:+.2e}")
        assert abs(rel) <= ENERGY_RTOL, s
    _say(f"[cfg] headline worst |rel| {worst:.2e}")


def test_config1_icosphere_pairs(gpu, oracle):
    _say("[cfg] config 1: icosphere L4 -> displaced L5, 64^3, 4 seeds")
    pairs = [synth.icosphere_pair(s, 4) for s in range(4)]
    assert all(len(p[0]) == 2562 and len(p[2]) == 10242 for p in pairs)
    outs, st = deform.rigid_deform_batch(pairs, grid_resolution=64, lambdas=1.0)
    assert all(s["solver_path_name"] == "lds" for s in st)
    for s, ((so, eg), gs) in enumerate(zip(_oracle_energies(oracle, pairs, 64, [1.0] * 4, outs), st)):
        rel = (eg - so["final_cost"]) / so["final_cost"]
        _say(f"[cfg] config 1 seed {s}: oracle E={so['final_cost']:.8g}  gpu E={eg:.8g} it={gs['iterations']} "
             f"rel {rel:+.2e}")
        assert abs(rel) <= ENERGY_RTOL, s


@pytest.fixture(scope="module")
def config2(gpu, oracle):
    """RigidLossLayer on the config-2 pair (5000-vertex box union, 128^3)."""
    from meshode_amd.layers import RigidLossLayer
    sV, sF, tV, tF = synth.box_union_pair(0)
    N = 128
    src = torch.from_numpy(sV.astype(np.float32))
    F = torch.from_numpy(sF.astype(np.int32))
    layer = RigidLossLayer(src, F, torch.from_numpy(tV.astype(np.float32)), torch.from_numpy(tF.astype(np.int32)),
                           grid_resolution=N)
    pid = int(layer.param_id)
    grid = pyDeform.TemplateGrid(pid)[0].cpu().numpy().reshape(-1)
    rest = src.numpy().copy()                 # normalized in place by the layer
    off = oracle.store_rigidity(rest, sF.astype(np.int32))
    return layer, pid, N, grid, rest, sF.astype(np.int32), off


def test_config2_layer_bit_exact(config2, oracle, gpu):
    layer, pid, N, grid, rest, F, off = config2
    _say("[cfg] config 2: RigidLossLayer at 128^3, outputs and gradient bit-exact")
    X0 = (rest + np.random.default_rng(5).normal(0, 0.004, rest.shape)).astype(np.float32)
    X = torch.from_numpy(X0).to(gpu).requires_grad_(True)
    Fd = torch.from_numpy(F).to(gpu)
    loss = layer(X, Fd)
    loss.backward()
    D = oracle.distance_loss_fwd(grid, N, X0)
    R = oracle.rigid_fwd(X0, F, off)
    assert np.array_equal(pyDeform.DistanceFieldLoss_forward(X.detach(), pid).cpu().numpy(), D)
    assert np.array_equal(pyDeform.RigidEdgeLoss_forward(X.detach(), Fd, pid).cpu().numpy(), R)
    want = 0.5 * D.astype(np.float64).sum() + 0.5 * R.astype(np.float64).sum()
    assert float(loss.detach()) == pytest.approx(want, rel=1e-5)
    g = oracle.distance_loss_bwd(grid, N, X0) + oracle.rigid_bwd(X0, F, off)
    assert np.array_equal(X.grad.cpu().numpy(), g)


def test_config2_gradient_vs_central_differences(config2, oracle, gpu):
    """FP64 central differences (h = 1e-4) of the FP64 loss terms touching
    each vertex (the FP64 lookup, uniformgrid.cc:18-83, on the FP64 grid;
    the FP64 edge residuals) against the GPU's FP32 gradient, on 64 vertices
    whose coordinates stay 2h N away from every cell face and whose distance
    is 1e-3 away from the 0.2 truncation.  Bar: 1e-4 relative to |FD| plus
    the sum of the magnitudes of the gradient's terms (the edge terms cancel
    near rest)."""
    layer, pid, N, grid, rest, F, off = config2
    _say("[cfg] config 2: gradient vs FP64 central differences on 64 vertices")
    rng = np.random.default_rng(6)
    X0 = (rest + rng.normal(0, 0.004, rest.shape)).astype(np.float32)
    X = torch.from_numpy(X0).to(gpu).requires_grad_(True)
    layer(X, torch.from_numpy(F).to(gpu)).backward()
    g = X.grad.cpu().numpy()
    gD = pyDeform.DistanceFieldLoss_backward(torch.from_numpy(X0).to(gpu), pid).cpu().numpy()
    h = 1e-4
    x64 = X0.astype(np.float64)
    d64 = oracle.lookup_f64(grid, N, x64, want_grad=False)
    fr = x64 * N - np.floor(x64 * N)
    ok = (np.all((fr > 2 * h * N) & (fr < 1 - 2 * h * N), 1) & (np.abs(d64 - 0.2) > 1e-3) &
          np.all(x64 * N < N - 2, 1) & np.all(x64 > 0, 1))
    pick = rng.choice(np.where(ok)[0], 64, replace=False)
    HE = oracle.face_edges(F)
    off64 = off.astype(np.float64)
    worst = 0.0
    for i in pick:
        hs = np.where((HE[:, 0] == i) | (HE[:, 1] == i))[0]

        def L(x):
            d = oracle.lookup_f64(grid, N, x[i:i + 1], want_grad=False)[0]
            r = x[HE[hs, 1]] - x[HE[hs, 0]] - off64[hs]
            return 0.5 * d * d + 0.5 * float((r * r).sum())

        r = x64[HE[hs, 1]] - x64[HE[hs, 0]] - off64[hs]
        for k in range(3):
            xp, xm = x64.copy(), x64.copy()
            xp[i, k] += h
            xm[i, k] -= h
            fd = (L(xp) - L(xm)) / (2 * h)
            tol = 1e-4 * (abs(fd) + abs(gD[i, k]) + np.abs(r[:, k]).sum())
            worst = max(worst, abs(g[i, k] - fd) / tol)
            assert abs(g[i, k] - fd) <= tol, (i, k, g[i, k], fd)
    _say(f"[cfg] config 2 FD: worst error / tolerance {worst:.3g}")


def test_config2_adam_steps(config2, oracle, gpu):
    """20 Adam steps at lr 1e-3 (rigid_deform.py:25-44): the GPU gradient
    equals the oracle's at every iterate, and the same loop on the CPU,
    driven by the oracle's layers, follows the same path."""
    layer, pid, N, grid, rest, F, off = config2
    _say("[cfg] config 2: 20 Adam steps, GPU layer vs oracle layers on the CPU")
    Fd = torch.from_numpy(F).to(gpu)
    Vg = torch.nn.Parameter(torch.from_numpy(rest.copy()).to(gpu))
    Vc = torch.nn.Parameter(torch.from_numpy(rest.copy()))
    og = torch.optim.Adam([Vg], lr=1e-3, foreach=False)
    oc = torch.optim.Adam([Vc], lr=1e-3, foreach=False)
    for step in range(20):
        og.zero_grad()
        loss = layer(Vg, Fd)
        loss.backward()
        x = Vg.detach().cpu().numpy()
        want = oracle.distance_loss_bwd(grid, N, x) + oracle.rigid_bwd(x, F, off)
        assert np.array_equal(Vg.grad.cpu().numpy(), want), step
        og.step()
        xc = Vc.detach().numpy()
        oc.zero_grad()
        Vc.grad = torch.from_numpy(oracle.distance_loss_bwd(grid, N, xc) + oracle.rigid_bwd(xc, F, off))
        oc.step()
    diff = np.abs(Vg.detach().cpu().numpy() - Vc.detach().numpy())
    moved = np.abs(Vc.detach().numpy() - rest).max()
    _say(f"[cfg] config 2 Adam: moved up to {moved:.3g}; GPU vs CPU loop max |dx| {diff.max():.3g}, "
         f"coordinates above 1e-6: {(diff > 1e-6).sum()}")
    assert moved > 5e-3
    assert (diff > 1e-6).mean() <= 1e-3
    assert diff.max() <= 2e-3


def test_config4_level6_pair_golden(gpu, oracle):
    """The largest config-4 pair: a 40962-vertex icosphere source against a
    displaced level-6 target at 128^3 (lambda 1.3), on the multi-workgroup
    register/LDS team kernel, within 1e-3 of the oracle's converged energy
    (tests/golden/lm_cfg4_l6_n128.npz, written by make_golden.py cfg4_l6)."""
    z = np.load(os.path.join(GOLDEN, "lm_cfg4_l6_n128.npz"))
    N, lam = int(z["N"]), float(z["lam"])
    _say(f"[cfg] config 4: level-6 pair at {N}^3 on the team kernel")
    pair = synth.icosphere_pair(int(z["seed"]), int(z["level"]))
    assert len(pair[0]) == 40962
    pid = deform.template_f64(pair[2], pair[3], N)
    _, scale, pos = pyDeform.TemplateInfo(pid)
    grid = pyDeform.TemplateGrid(pid)[0].cpu().numpy().reshape(-1)
    sn = oracle.apply_transform(pair[0], scale, pos)
    V = torch.from_numpy(sn.copy()).cuda()
    st = deform.deform_batch([pid], [V], [pair[1]], [0], [lam])[0]
    assert st["solver_path_name"] == "lds_team"
    eg, _, _ = oracle.lm_cost(grid, N, V.cpu().numpy(), sn, oracle.face_edges(pair[1]), lam)
    ref = float(z["final_cost"])
    assert float(z["initial_cost"]) == pytest.approx(st["initial_cost"], rel=1e-4)
    _say(f"[cfg] level 6: oracle E={ref:.8g} it={int(z['iterations'])}  gpu E={eg:.8g} it={st['iterations']}")
    assert abs(eg - ref) <= ENERGY_RTOL * ref


def test_config4_mixed_batch_alone_equals_mix(gpu):
    """Eight config-4 pairs (levels 3-6, 64^3 / 128^3, lambda in [0.5, 2],
    50 LM x 30 PCG iterations): the one-workgroup and the team register/LDS
    kernels in one call; each pair's result is bit-identical when solved
    alone."""
    _say("[cfg] config 4: mixed L3-L6 batch, alone vs in the mix")
    rng = np.random.Generator(np.random.PCG64(4))
    levels = [3, 4, 5, 6, 3, 4, 5, 6]
    grids = [int(g) for g in rng.choice([64, 128], 8)]
    lams = [float(x) for x in rng.uniform(0.5, 2.0, 8)]
    pairs = [synth.icosphere_pair(1000 * L + i, L) for i, L in enumerate(levels)]
    opt = deform.default_options(max_num_iterations=50, pcg_max_iterations=30, function_tolerance=0.0,
                                 gradient_tolerance=0.0, parameter_tolerance=0.0, pcg_relative_tolerance=0.0)
    mix, st_mix = deform.rigid_deform_batch(pairs, grid_resolution=grids, lambdas=lams, options=opt)
    paths = {s["solver_path_name"] for s in st_mix}
    assert {"lds", "lds_team"} <= paths, paths
    for i, p in enumerate(pairs):
        alone, st = deform.rigid_deform_batch([p], grid_resolution=grids[i], lambdas=lams[i], options=opt)
        assert torch.equal(alone[0], mix[i]), i
        assert st[0]["final_cost"] == st_mix[i]["final_cost"], i
        assert st[0]["iterations"] <= 50
