"""GPU parity of the NeuralODE bidirectional path (cad_neural_deform2.py)
against the oracle.

* SolveLinear(with_rot = 1) / mo_linear_estimation_rot (linear.cc:114-230)
  against the oracle's refined solve, on an icosphere and on subdivided CAD
  meshes (one with coincident and 5.6e-17-apart vertices joined by face
  edges), FP64 and through the float32 pyDeform binding; bad indices are
  errors.
* mo_nearest_f32 (ReverseLossLayer's k = 1 query) bit-exact (index and FP64
  squared distance) against the oracle's brute force, with duplicates and
  exact ties, over launch shapes with one and many source chunks, and on
  the uniform-grid search (4096 sources or more): box faces with near, far
  and inside targets, zero extent along an axis or all axes, duplicated
  sources and lattice ties.
* ReverseLossLayer loss and source gradient; GraphLoss2Layer forward and
  backward against the oracle's distance and graph-edge losses (bit-exact
  gradient) with each graph normalized by its own template.
* The driver: the first iterations of the bidirectional Adam loop against
  the same loop on the CPU with the oracle's losses and nearest neighbours
  (same initial weights), and the final SolveLinear(with_rot = 1) against
  the oracle on the driver's own pushed positions.
"""
import numpy as np
import pytest
import torch
from torch.autograd import Function

from meshode_amd import cad, objio, pyDeform, synth
from meshode_amd.layers import graph_loss2_layer, reverse_loss_layer
from meshode_amd.layers import neuralode_fast as nf
from oracle import cad_oracle, neural_oracle

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, flush=True)


def _smooth_target(V, seed, amp=0.03):
    rng = np.random.default_rng(seed)
    a = rng.normal(0, 1, (3, 3))
    return V + amp * np.sin(3.0 * V @ a)


def _cad_mesh(extent, tmp_path=None, center=(0.5, 0.5, 0.5)):
    """A subdivided CAD box; read back from an OBJ file when tmp_path is
    given (those coordinates leave coincident and 5.6e-17-apart vertices
    joined by face edges)."""
    V, F = synth.box_surface(center, extent, (1, 1, 1))
    if tmp_path is not None:
        objio.write_obj(str(tmp_path / "box.obj"), V, F)
        V, F = objio.read_obj(str(tmp_path / "box.obj"))
    return cad.prepare(V, F)


@pytest.mark.parametrize("case", ["icosphere_l3", "cad_box", "cad_box_obj"])
def test_linear_estimation_rot_against_oracle(gpu, case, tmp_path):
    _say(f"[neural] LinearEstimationWithRot {case}")
    if case == "icosphere_l3":
        V, F = synth.icosphere(3)
        V = 0.5 + 0.3 * V
    else:
        r = _cad_mesh((0.2, 0.14, 0.1), tmp_path if case == "cad_box_obj" else None)
        V, F = r["V"], r["F"]
    TV = _smooth_target(V, 1)
    X = torch.from_numpy(V).to(gpu).contiguous()
    it = cad.linear_estimation_rot(X, torch.from_numpy(F.astype(np.int32)).to(gpu), torch.from_numpy(TV).to(gpu), 1.0)
    got = X.cpu().numpy()
    want = cad_oracle.linear_estimation_rot(V, F, TV, 1.0)
    _say(f"[neural]   n={len(V)} CG iterations {it} max |diff| {np.abs(got - want).max():.3e}")
    np.testing.assert_allclose(got, want, rtol=0, atol=1e-10)
    # the float32 binding (the reference's tensors): inputs read as float32
    Vt = torch.from_numpy(V.astype(np.float32))
    TVt = torch.from_numpy(TV.astype(np.float32))
    pyDeform.SolveLinear(Vt, torch.from_numpy(F.astype(np.int32)), torch.zeros(0, 2, dtype=torch.int32),
                         torch.arange(len(V), dtype=torch.int32), TVt, 1, 1)
    want32 = cad_oracle.linear_estimation_rot(V.astype(np.float32).astype(np.float64), F,
                                              TV.astype(np.float32).astype(np.float64), 1.0)
    assert Vt.device.type == "cpu"
    np.testing.assert_allclose(Vt.numpy(), want32.astype(np.float32), rtol=0, atol=2e-7)


def test_linear_estimation_rot_errors(gpu):
    _say("[neural] LinearEstimationWithRot coincident pair and error paths")
    from meshode_amd import _lib
    V = torch.rand(4, 3, dtype=torch.float64, device=gpu)
    F = torch.tensor([[0, 1, 2], [0, 2, 3]], dtype=torch.int32, device=gpu)
    V[3] = V[2]                               # zero-length edge (2, 3): one unknown
    X = V.clone()
    cad.linear_estimation_rot(X, F, V.clone() + 0.01, 1.0)
    assert torch.equal(X[2], X[3])
    np.testing.assert_allclose(X.cpu().numpy(), cad_oracle.linear_estimation_rot(
        V.cpu().numpy(), F.cpu().numpy(), V.cpu().numpy() + 0.01, 1.0), rtol=0, atol=1e-12)
    with pytest.raises(_lib.MoError, match="out of range"):
        cad.linear_estimation_rot(torch.rand(4, 3, dtype=torch.float64, device=gpu),
                                  torch.tensor([[0, 1, 4]], dtype=torch.int32, device=gpu),
                                  torch.rand(4, 3, dtype=torch.float64, device=gpu), 1.0)
    with pytest.raises(ValueError):
        cad.linear_estimation_rot(V.float(), F, V, 1.0)


@pytest.mark.parametrize("ns,nt", [(1, 5), (1000, 7), (5000, 3000), (70001, 100), (3, 0)])
def test_nearest_bit_exact(gpu, ns, nt):
    _say(f"[neural] nearest ns={ns} nt={nt}")
    rng = np.random.default_rng(ns + nt)
    src = rng.random((ns, 3)).astype(np.float32)
    tar = rng.random((nt, 3)).astype(np.float32)
    if ns > 10 and nt > 4:
        src[ns // 2] = src[3]                 # duplicate source points: the lower index wins
        tar[:2] = src[3]
        tar[2] = src[ns - 1]                  # exact hit on the last source point
        tar[3] = (src[5] + src[6]) / 2        # near the midpoint of two points
    idx, d2 = reverse_loss_layer.nearest(torch.from_numpy(src).to(gpu), torch.from_numpy(tar).to(gpu), True)
    k, od2 = neural_oracle.nearest(src, tar)
    assert np.array_equal(idx.cpu().numpy(), k)
    assert np.array_equal(d2.cpu().numpy(), od2)
    if ns > 10 and nt > 4:
        assert k[0] == 3 and k[2] == ns - 1 and od2[2] == 0.0


def _nn_case(name, rng):
    """Source / target sets for the grid search (4096 sources or more)."""
    if name == "box_surface":      # a box's faces, targets near and far, one inside
        n = 6000
        u = rng.random((n, 2)).astype(np.float32)
        face = rng.integers(0, 6, n)
        P = np.zeros((n, 3), np.float32)
        for f in range(6):
            m = face == f
            ax = f // 2
            oth = [a for a in range(3) if a != ax]
            P[m, ax] = 0.2 + 0.6 * (f % 2)
            P[m, oth[0]] = 0.2 + 0.6 * u[m, 0]
            P[m, oth[1]] = 0.2 + 0.6 * u[m, 1]
        tar = (P[rng.integers(0, n, 3000)] + rng.normal(0, 0.02, (3000, 3))).astype(np.float32)
        return P, np.vstack([tar, np.array([[5, 5, 5], [-3, 0.5, 0.5], [0.5, 0.5, 0.5]], np.float32)])
    if name == "flat_z":           # zero extent along z
        P = rng.random((4500, 3)).astype(np.float32)
        P[:, 2] = 0.25
        return P, rng.random((500, 3)).astype(np.float32)
    if name == "all_equal":        # zero extent: one cell
        return np.tile(np.array([[0.3, 0.4, 0.5]], np.float32), (4200, 1)), rng.random((60, 3)).astype(np.float32)
    if name == "line_dups":        # duplicated sources: the lower index wins
        line = np.zeros((4100, 3), np.float32)
        line[:, 0] = np.linspace(0, 1, 4100)
        line = np.vstack([line, line[:50]])
        return line, np.vstack([line[:20] + 1e-7, rng.random((40, 3)).astype(np.float32)])
    if name == "lattice_ties":     # targets at cell centres: up to 8 sources equally near
        g = np.stack(np.meshgrid(*[np.arange(17)] * 3, indexing="ij"), -1).reshape(-1, 3).astype(np.float32) / 16
        return g, (g[rng.integers(0, len(g), 2000)] + 1 / 32).astype(np.float32)
    raise KeyError(name)


@pytest.mark.parametrize("name", ["box_surface", "flat_z", "all_equal", "line_dups", "lattice_ties"])
def test_nearest_grid_search_bit_exact(gpu, name):
    """The uniform-grid search (4096 sources or more) returns the brute
    force's index and FP64 distance bits, ties to the lowest index."""
    _say(f"[neural] nearest grid search {name}")
    src, tar = _nn_case(name, np.random.default_rng(11))
    idx, d2 = reverse_loss_layer.nearest(torch.from_numpy(src).to(gpu), torch.from_numpy(tar).to(gpu), True)
    k, od2 = neural_oracle.nearest(src, tar)
    assert np.array_equal(idx.cpu().numpy(), k)
    assert np.array_equal(d2.cpu().numpy(), od2)


def test_reverse_loss_layer(gpu):
    _say("[neural] ReverseLossLayer loss and gradient")
    rng = np.random.default_rng(11)
    src = rng.random((1500, 3)).astype(np.float32)
    tar = rng.random((900, 3)).astype(np.float32)
    S = torch.from_numpy(src).to(gpu).requires_grad_(True)
    loss = reverse_loss_layer.ReverseLossLayer()(S, torch.from_numpy(tar), gpu)
    loss.backward()
    k, _ = neural_oracle.nearest(src, tar)
    assert float(loss.detach()) == pytest.approx(neural_oracle.reverse_loss(src, tar), rel=1e-5)
    g = np.zeros_like(src, dtype=np.float64)
    np.add.at(g, k, src[k].astype(np.float64) - tar)
    np.testing.assert_allclose(S.grad.cpu().numpy(), g, rtol=0, atol=1e-5)


def _load_pair(tmp_path):
    """Two CAD boxes through LoadCadMesh (the driver's input bundles)."""
    out = []
    for name, ext, ctr in (("src", (0.2, 0.14, 0.1), (0.5, 0.5, 0.5)), ("tgt", (0.17, 0.16, 0.12), (0.52, 0.49, 0.5))):
        V, F = synth.box_surface(ctr, ext, (1, 1, 1))
        p = str(tmp_path / f"{name}.obj")
        objio.write_obj(p, V, F)
        out.append(pyDeform.LoadCadMesh(p))
    return out


class _OracleGraphLoss2(Function):
    """GraphLoss2Function on the CPU with the oracle's losses."""

    @staticmethod
    def forward(ctx, V, O, grid, N, E, off, r2):
        v = V.detach().numpy()
        D = O.distance_loss_fwd(grid, N, v) * np.float32(0.5)
        R = O.graph_fwd(v, E, off) * np.float32(0.5)
        ctx.save_for_backward(V)
        ctx.args = (O, grid, N, E, off, r2)
        return torch.from_numpy(D).sum() + torch.from_numpy(R).sum() * r2

    @staticmethod
    def backward(ctx, g):
        (V,) = ctx.saved_tensors
        O, grid, N, E, off, r2 = ctx.args
        v = V.detach().numpy()
        gD = torch.from_numpy(O.distance_loss_bwd(grid, N, v))
        gR = torch.from_numpy(O.graph_bwd(v, E, off))
        return g * (gD + gR * r2), None, None, None, None, None, None


def test_graph_loss2_layer(gpu, oracle, tmp_path):
    _say("[neural] GraphLoss2Layer forward/backward")
    (V1, F1, E1, R1, GV1, GE1), (V2, F2, E2, R2, GV2, GE2) = _load_pair(tmp_path)
    GV1_raw, GV2_raw = GV1.clone(), GV2.clone()
    layer = graph_loss2_layer.GraphLoss2Layer(V1, F1, GV1, GE1, V2, F2, GV2, GE2, 0.1, gpu)
    p1, p2 = int(layer.param_id1), int(layer.param_id2)
    N1, s1, t1 = pyDeform.TemplateInfo(p1)
    N2, s2, t2 = pyDeform.TemplateInfo(p2)
    # each graph normalized in place by its own template
    assert np.array_equal(GV1.numpy(), oracle.normalize_by_template_f32(GV1_raw.numpy(), s1, t1))
    assert np.array_equal(GV2.numpy(), oracle.normalize_by_template_f32(GV2_raw.numpy(), s2, t2))
    grid1 = pyDeform.TemplateGrid(p1)[0].cpu().numpy().reshape(-1)
    grid2 = pyDeform.TemplateGrid(p2)[0].cpu().numpy().reshape(-1)
    off1 = oracle.store_graph(GV1.numpy(), GE1.numpy())
    off2 = oracle.store_graph(GV2.numpy(), GE2.numpy())
    r2 = torch.tensor(0.1 * 0.1).tolist()
    rng = np.random.default_rng(5)
    for direction in (0, 1):
        base = GV1 if direction == 0 else GV2
        P = (base.numpy() + rng.normal(0, 0.01, base.shape)).astype(np.float32)
        X = torch.from_numpy(P).to(gpu).requires_grad_(True)
        if direction == 0:
            loss = layer(X, GE1.to(gpu), GV2.to(gpu), GE2.to(gpu), 0)
            grid, N, E, off = grid2, N2, GE1.numpy(), off1
        else:
            loss = layer(GV1.to(gpu), GE1.to(gpu), X, GE2.to(gpu), 1)
            grid, N, E, off = grid1, N1, GE2.numpy(), off2
        loss.backward()
        D = oracle.distance_loss_fwd(grid, N, P) * np.float32(0.5)
        R = oracle.graph_fwd(P, E, off) * np.float32(0.5)
        want = float(D.astype(np.float64).sum() + R.astype(np.float64).sum() * np.float32(r2))
        assert float(loss) == pytest.approx(want, rel=1e-5)
        gD = oracle.distance_loss_bwd(grid, N, P)
        gR = oracle.graph_bwd(P, E, off)
        # torch's float32 arithmetic: gD + gR * float32(r2)
        assert np.array_equal(X.grad.cpu().numpy(), gD + gR * np.float32(r2))


def test_cad_neural_deform2_driver(gpu, oracle, tmp_path):
    from meshode_amd import cad_neural_deform2 as drv
    _say("[neural] cad_neural_deform2 driver, 4 iterations against the CPU loop")
    m1, m2 = _load_pair(tmp_path)
    V1_raw = m1[0].clone()
    GV1_raw, GV2_raw = m1[4].clone(), m2[4].clone()
    niter = 4
    torch.manual_seed(1234)
    r = drv.deform(m1, m2, 0.1, gpu, niter, log=_say)
    p1, p2 = int(r["param_id1"]), int(r["param_id2"])
    N1, s1, t1 = pyDeform.TemplateInfo(p1)
    N2, s2, t2 = pyDeform.TemplateInfo(p2)
    grid1 = pyDeform.TemplateGrid(p1)[0].cpu().numpy().reshape(-1)
    grid2 = pyDeform.TemplateGrid(p2)[0].cpu().numpy().reshape(-1)
    # the same loop on the CPU: oracle losses, oracle nearest neighbours
    G1 = torch.from_numpy(oracle.normalize_by_template_f32(GV1_raw.numpy(), s1, t1))
    G2 = torch.from_numpy(oracle.normalize_by_template_f32(GV2_raw.numpy(), s2, t2))
    E1n, E2n = m1[5].numpy(), m2[5].numpy()
    off1, off2 = oracle.store_graph(G1.numpy(), E1n), oracle.store_graph(G2.numpy(), E2n)
    r2 = torch.tensor(0.1 * 0.1).tolist()
    torch.manual_seed(1234)
    func = nf.NeuralODE(torch.device("cpu"))
    opt = torch.optim.Adam(func.parameters(), lr=1e-3)

    def rev(src, tar):
        k, _ = neural_oracle.nearest(src.detach().numpy(), tar.numpy())
        d = src[torch.from_numpy(k)] - tar
        return 0.5 * (d * d).sum()

    cpu = []
    for _ in range(niter):
        opt.zero_grad()
        a = func.forward(G1)
        b = func.inverse(G2)
        l1f = _OracleGraphLoss2.apply(a, oracle, grid2, N2, E1n, off1, r2)
        l1b = rev(a, G2)
        l2f = _OracleGraphLoss2.apply(b, oracle, grid1, N1, E2n, off2, r2)
        l2b = rev(b, G1)
        (l1f + l1b + l2f + l2b).backward()
        opt.step()
        cpu.append([float(l1f), float(l1b), float(l2f), float(l2b)])
    cpu = np.array(cpu)
    _say(f"[neural]   GPU losses {r['losses'][-1]}, CPU {cpu[-1]}")
    np.testing.assert_allclose(r["losses"][0], cpu[0], rtol=1e-5)
    np.testing.assert_allclose(r["losses"], cpu, rtol=1e-3)
    # final LinearEstimationWithRot on the driver's own pushed positions
    V1n = oracle.normalize_by_template_f32(V1_raw.numpy(), s1, t1).astype(np.float64)
    TV = r["V1_pushed"].cpu().numpy().astype(np.float64)
    X = cad_oracle.linear_estimation_rot(V1n, m1[1].numpy(), TV, 1.0).astype(np.float32)
    want = oracle.denormalize_by_template_f32(X, s2, t2)
    got = r["V"].numpy()
    assert got.shape == want.shape
    np.testing.assert_allclose(got, want, rtol=0, atol=4 * np.spacing(np.abs(want).max()))
