"""GPU parity of the pyDeform layer functions against the oracle, bit-exact:
NormalizeByTemplate / DenormalizeByTemplate (normalize.cc:5-45),
DistanceFieldLoss_forward/backward (distance_layer.cc:8-80),
Store*/RigidEdgeLoss_*/GraphEdgeLoss_* (rigid_layer.cc, graph_layer.cc)."""
import numpy as np
import pytest
import torch

from meshode_amd import pyDeform, synth

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, flush=True)


@pytest.fixture(scope="module")
def sphere_template(gpu, oracle):
    _say("[layers] template sphere L3 N=32")
    _, _, V, F = synth.displaced_sphere_pair(7, src_level=2, tgt_level=3)
    Vt = torch.from_numpy(V.astype(np.float32))
    Ft = torch.from_numpy(F.astype(np.int32))
    N = 32
    pid = pyDeform.InitializeDeformTemplate(Vt, Ft, 0, N)
    d64, _, _ = pyDeform.TemplateGrid(pid)
    return pid, N, d64.cpu().numpy().reshape(-1)


def _probe_points(N, n, seed):
    rng = np.random.default_rng(seed)
    pts = [rng.uniform(0.0, 1.0, (n, 3))]
    # exact cell corners and faces (index boundaries), the N-1 band, outside the box
    k = rng.integers(0, N + 1, (n // 4, 3))
    pts.append(k / N)
    pts.append(rng.uniform((N - 1) / N, 1.0, (n // 8, 3)))
    pts.append(rng.uniform(-0.3, 1.3, (n // 4, 3)))
    pts.append(rng.uniform(-1.0 / N, 0.0, (n // 8, 3)))
    P = np.concatenate(pts).astype(np.float32)
    return P


@pytest.mark.parametrize("device", ["cpu", "cuda"])
def test_distance_loss_bit_exact(sphere_template, oracle, device):
    pid, N, grid = sphere_template
    _say(f"[layers] distance loss fwd/bwd tensors on {device}")
    P = _probe_points(N, 20000, 1)
    Vt = torch.from_numpy(P).to(device)
    fwd = pyDeform.DistanceFieldLoss_forward(Vt, pid)
    bwd = pyDeform.DistanceFieldLoss_backward(Vt, pid)
    assert fwd.device.type == device and bwd.device.type == device
    assert fwd.shape == (len(P),) and bwd.shape == (len(P), 3)
    of = oracle.distance_loss_fwd(grid, N, P)
    ob = oracle.distance_loss_bwd(grid, N, P)
    assert np.array_equal(fwd.cpu().numpy(), of)
    assert np.array_equal(bwd.cpu().numpy(), ob)
    # truncated, zero-penalty and penalty branches all appear in the probe set
    d = oracle.lookup_f32(grid, N, P, want_grad=False)
    assert (d == 0).any() and (d > 0.2).any()


def test_normalize_by_template_bit_exact(sphere_template, oracle):
    pid, N, _ = sphere_template
    _say("[layers] normalize/denormalize by template")
    _, scale, trans = pyDeform.TemplateInfo(pid)
    rng = np.random.default_rng(2)
    V = rng.normal(0, 1, (5000, 3)).astype(np.float32)
    Vt = torch.from_numpy(V.copy()).cuda()
    pyDeform.NormalizeByTemplate(Vt, pid)
    exp = oracle.normalize_by_template_f32(V, scale, trans)
    assert np.array_equal(Vt.cpu().numpy(), exp)
    pyDeform.DenormalizeByTemplate(Vt, pid)
    exp2 = oracle.denormalize_by_template_f32(exp, scale, trans)
    assert np.array_equal(Vt.cpu().numpy(), exp2)
    # CPU tensor, updated in place
    Vc = torch.from_numpy(V.copy())
    pyDeform.NormalizeByTemplate(Vc, pid)
    assert np.array_equal(Vc.numpy(), exp)


def test_rigid_loss_bit_exact(sphere_template, oracle):
    pid, N, _ = sphere_template
    _say("[layers] rigid edge loss")
    sV, sF, _, _ = synth.displaced_sphere_pair(7, src_level=3, tgt_level=3)
    V0 = (sV * 0.3 + 0.5).astype(np.float32)
    F = sF.astype(np.int32)
    Vt, Ft = torch.from_numpy(V0), torch.from_numpy(F)
    pyDeform.StoreRigidityInformation(Vt, Ft, pid)
    rng = np.random.default_rng(3)
    V1 = (V0 + rng.normal(0, 0.01, V0.shape)).astype(np.float32)
    V1t = torch.from_numpy(V1).cuda()
    fwd = pyDeform.RigidEdgeLoss_forward(V1t, Ft.cuda(), pid).cpu().numpy()
    bwd = pyDeform.RigidEdgeLoss_backward(V1t, Ft.cuda(), pid).cpu().numpy()
    off = oracle.store_rigidity(V0, F)
    assert np.array_equal(fwd, oracle.rigid_fwd(V1, F, off))
    assert np.array_equal(bwd, oracle.rigid_bwd(V1, F, off))
    # zero at rest
    assert float(pyDeform.RigidEdgeLoss_forward(Vt, Ft, pid).abs().max()) == 0.0


def test_graph_loss_bit_exact(sphere_template, oracle):
    pid, N, _ = sphere_template
    _say("[layers] graph edge loss")
    rng = np.random.default_rng(4)
    nv = 3000
    V0 = rng.uniform(0.2, 0.8, (nv, 3)).astype(np.float32)
    E = rng.integers(0, nv, (12000, 2)).astype(np.int32)
    E[:50, 1] = E[:50, 0]                  # degenerate edges (a == b)
    E[100:200] = E[200:300]                # duplicated edges
    Vt, Et = torch.from_numpy(V0), torch.from_numpy(E)
    pyDeform.StoreGraphInformation(Vt, Et, pid)
    V1 = (V0 + rng.normal(0, 0.01, V0.shape)).astype(np.float32)
    V1t = torch.from_numpy(V1)
    fwd = pyDeform.GraphEdgeLoss_forward(V1t, Et, pid).numpy()
    bwd = pyDeform.GraphEdgeLoss_backward(V1t, Et, pid).numpy()
    off = oracle.store_graph(V0, E)
    assert np.array_equal(fwd, oracle.graph_fwd(V1, E, off))
    assert np.array_equal(bwd, oracle.graph_bwd(V1, E, off))


def test_layer_errors(sphere_template):
    pid, N, _ = sphere_template
    _say("[layers] error paths")
    from meshode_amd import _lib
    V = torch.rand(10, 3)
    with pytest.raises(_lib.MoError):
        pyDeform.DistanceFieldLoss_forward(V, 10_000)
    with pytest.raises(TypeError):
        pyDeform.DistanceFieldLoss_forward(V.double(), pid)
    Fbad = torch.tensor([[0, 1, 99]], dtype=torch.int32)
    with pytest.raises(_lib.MoError):
        pyDeform.StoreRigidityInformation(V, Fbad, pid)
    # a fresh template has no stored offsets
    Vt = torch.rand(20, 3)
    Ft = torch.randint(0, 20, (30, 3), dtype=torch.int32)
    pid2 = pyDeform.InitializeDeformTemplate(Vt, Ft, 0, 8)
    with pytest.raises(_lib.MoError):
        pyDeform.RigidEdgeLoss_forward(Vt, Ft, pid2)
