"""Oracle pins for the NeuralODE bidirectional path (cad_neural_deform2.py).

* LinearEstimationWithRot (linear.cc:114-230): the frames reduce to a closed
  form when the target is a similarity of the rest mesh (R = the rotation,
  scale = s sum|d1| / (sum|d1| + 1e-8)); for a general target R is the
  orthogonal polar factor (R^T C symmetric positive semidefinite); the
  assembled solve equals a least-squares solve of the energy written term by
  term (one row per vertex pull and per half-edge direction), which fails
  for a dropped term, a wrong sign or a swapped direction in the assembly.
* ReverseLossLayer's nearest neighbours: the oracle's brute force against
  scipy's cKDTree (the reference's own call) where the nearest point is
  unique by a margin; exact ties go to the lowest index.
* The fixed-grid 3/8-rule odeint (torchdiffeq is absent, so its stepping is
  pinned by mathematics): exact for y' = cubic in t (the 3/8 quadrature
  rule), equal to the degree-4 Taylor polynomial of exp(A) on y' = A y,
  and the decreasing grid [1, 0] equals the step with dt = -1.
"""
import math

import numpy as np
import pytest
import torch
from scipy.spatial import cKDTree
from scipy.spatial.transform import Rotation

from meshode_amd import synth
from meshode_amd.layers import neuralode_fast as nf
from oracle import cad_oracle, neural_oracle


def _mesh(level=1, seed=0):
    V, F = synth.icosphere(level)
    rng = np.random.default_rng(seed)
    V = 0.5 + 0.3 * V * (1 + 0.05 * rng.standard_normal((len(V), 1)))
    return V, F.astype(np.int64)


def _neighbour_sums(V, F):
    n = len(V)
    links = [set() for _ in range(n)]
    for f in F:
        for j in range(3):
            a, b = int(f[j]), int(f[(j + 1) % 3])
            links[a].add(b)
            links[b].add(a)
    return np.array([sum(np.linalg.norm(V[p] - V[i]) for p in sorted(links[i])) for i in range(n)])


def test_frames_of_a_similarity():
    V, F = _mesh(1)
    Rm = Rotation.random(random_state=3).as_matrix()
    s = 1.7
    TV = s * V @ Rm.T + np.array([0.1, -0.2, 0.05])
    scale, R = cad_oracle.rot_frames(V, F, TV)
    lo = _neighbour_sums(V, F)
    np.testing.assert_allclose(R, np.broadcast_to(Rm, R.shape), rtol=0, atol=1e-13)
    np.testing.assert_allclose(scale, s * lo / (lo + 1e-8), rtol=1e-13)


def test_frames_are_polar_factors():
    V, F = _mesh(1, seed=1)
    rng = np.random.default_rng(2)
    TV = V + 0.05 * rng.standard_normal(V.shape)
    scale, R = cad_oracle.rot_frames(V, F, TV)
    n = len(V)
    links = [set() for _ in range(n)]
    for f in F:
        for j in range(3):
            links[f[j]].add(f[(j + 1) % 3])
            links[f[(j + 1) % 3]].add(f[j])
    for i in range(n):
        C = sum(np.outer(TV[p] - TV[i], V[p] - V[i]) for p in links[i])
        np.testing.assert_allclose(R[i] @ R[i].T, np.eye(3), atol=1e-12)
        P = R[i].T @ C
        np.testing.assert_allclose(P, P.T, atol=1e-12 * np.abs(C).max())
        assert np.linalg.eigvalsh((P + P.T) / 2).min() > -1e-12 * np.abs(C).max()


def _lstsq_rot(V, F, TV, rigidity, scale, R):
    """The energy written as weighted least-squares rows, solved densely."""
    n = len(V)
    rows, rhs = [], []
    for i in range(n):
        r = np.zeros(n)
        r[i] = 1.0
        rows.append(r)
        rhs.append(TV[i])
    for f in F:
        for k in range(3):
            e = (int(f[k]), int(f[(k + 1) % 3]))
            for j in range(2):
                v0, v1 = e[j], e[1 - j]
                off = V[v1] - V[v0]
                w = math.sqrt(2e-2 / np.linalg.norm(off) * rigidity)
                r = np.zeros(n)
                r[v1] += w
                r[v0] -= w
                rows.append(r)
                rhs.append(w * scale[v0] * (R[v0] @ off))
    X, *_ = np.linalg.lstsq(np.array(rows), np.array(rhs), rcond=None)
    return X


@pytest.mark.parametrize("rigidity", [1.0, 0.3])
def test_rot_solve_equals_least_squares(rigidity):
    V, F = _mesh(1, seed=3)
    rng = np.random.default_rng(4)
    TV = V + 0.03 * rng.standard_normal(V.shape)
    X = cad_oracle.linear_estimation_rot(V, F, TV, rigidity)
    scale, R = cad_oracle.rot_frames(V, F, TV)
    np.testing.assert_allclose(X, _lstsq_rot(V, F, TV, rigidity, scale, R), rtol=0, atol=1e-12)


def test_rot_solve_similarity_fixed_point():
    """A similarity target is a fixed point up to the +1e-8 in the scale's
    denominator (relative 1e-8 / sum|d1|)."""
    V, F = _mesh(2, seed=5)
    Rm = Rotation.random(random_state=6).as_matrix()
    TV = 1.2 * V @ Rm.T + 0.1
    X = cad_oracle.linear_estimation_rot(V, F, TV, 1.0)
    lo = _neighbour_sums(V, F)
    assert np.abs(X - TV).max() < 10 * 1e-8 / lo.min() * np.abs(TV - TV.mean(0)).max()


def test_nearest_against_ckdtree():
    rng = np.random.default_rng(7)
    src = rng.random((3000, 3)).astype(np.float32)
    tar = rng.random((2000, 3)).astype(np.float32)
    k, d2 = neural_oracle.nearest(src, tar)
    dd, ii = cKDTree(src.astype(np.float64)).query(tar.astype(np.float64), k=2)
    unique = (dd[:, 1] - dd[:, 0]) > 1e-9
    assert unique.mean() > 0.99
    assert np.array_equal(k[unique], ii[unique, 0])
    np.testing.assert_allclose(np.sqrt(d2), dd[:, 0], rtol=1e-12)


def test_nearest_ties_lowest_index():
    src = np.array([[0, 0, 0], [1, 0, 0], [1, 0, 0], [0, 2, 0]], np.float32)
    tar = np.array([[0.5, 0, 0], [1, 0, 0], [0, 1, 0], [0.9, 0, 0]], np.float32)
    k, d2 = neural_oracle.nearest(src, tar)
    assert k.tolist() == [0, 1, 0, 1]
    assert d2[1] == 0.0


def test_odeint_cubic_in_time_is_exact():
    """y' = 1 + t + t^2 + t^3: the 3/8 rule integrates cubics exactly."""
    f = lambda t, y: (1 + t + t * t + t * t * t) * torch.ones_like(y)     # noqa: E731
    y0 = torch.zeros(2, 3, dtype=torch.float64)
    t = torch.tensor([0.0, 1.0], dtype=torch.float64)
    y1 = nf.odeint(f, y0, t)[1]
    assert torch.allclose(y1, torch.full_like(y0, 1 + 1 / 2 + 1 / 3 + 1 / 4), rtol=0, atol=1e-15)
    # backwards over [1, 0] returns to 0
    assert torch.allclose(nf.odeint(f, y1, torch.tensor([1.0, 0.0], dtype=torch.float64))[1], y0, atol=1e-15)


def test_odeint_linear_taylor_both_directions():
    torch.manual_seed(0)
    A = torch.randn(3, 3, dtype=torch.float64) * 0.4
    f = lambda t, y: y @ A.t()      # noqa: E731
    y = torch.randn(4, 3, dtype=torch.float64)
    for t, h in (([0.0, 1.0], 1.0), ([1.0, 0.0], -1.0), ([0.0, 0.5, 1.0], 0.5)):
        M = h * A
        P = torch.eye(3, dtype=torch.float64) + M + M @ M / 2 + M @ M @ M / 6 + M @ M @ M @ M / 24
        steps = len(t) - 1
        expect = y @ torch.linalg.matrix_power(P, steps).t()
        got = nf.odeint(f, y, torch.tensor(t, dtype=torch.float64))[-1]
        assert torch.allclose(got, expect, rtol=1e-14, atol=1e-14)


def test_odefunc_structure_and_time_input():
    torch.manual_seed(1)
    func = nf.ODEFunc().double()
    shapes = [tuple(p.shape) for p in func.parameters()]
    assert shapes == [(50, 4), (50,), (50, 50), (50,), (50, 50), (50,), (3, 50), (3,)]
    assert all(float(b.detach().abs().max()) == 0 for b in list(func.parameters())[1::2])
    y = torch.rand(6, 3, dtype=torch.float64)
    t = torch.tensor(0.25, dtype=torch.float64)
    # the time enters as a fourth coordinate, and both are centred by 0.5
    W0 = func.net[0].weight
    h = torch.nn.functional.leaky_relu((y - 0.5) @ W0[:, :3].t() + (0.25 - 0.5) * W0[:, 3])
    for lin in (func.net[2], func.net[4]):
        h = torch.nn.functional.leaky_relu(h @ lin.weight.t())
    assert torch.allclose(func(t, y), h @ func.net[6].weight.t(), rtol=1e-13, atol=1e-15)


def test_rot_solve_coincident_vertices_limit(tmp_path):
    """Subdivide leaves coincident vertices joined by face edges on
    axis-aligned faces (a lattice point on a split edge); the reference's
    weight 2e-2/|e| is infinite there.  Pulling each coincident copy apart by
    eps (a finite system of the reference's form) converges to the merged
    solution at rate eps (vertices 5.6e-17 apart, which the same meshes also
    hold, need the refined solve for this: a direct LU of the assembled
    matrix is off by ~2e-2 there)."""
    from meshode_amd import cad, objio
    V0, F0 = synth.box_surface((0.5, 0.5, 0.5), (0.2, 0.14, 0.1), (1, 1, 1))
    objio.write_obj(str(tmp_path / "box.obj"), V0, F0)          # the coordinates an OBJ file holds
    r = cad.prepare(*objio.read_obj(str(tmp_path / "box.obj")))
    V, F = r["V"], r["F"]
    cls = cad_oracle.coincident_classes(V, F)
    assert cls.max() + 1 < len(V)                     # the case occurs
    field = lambda P: P + 0.03 * np.sin(3.0 * P @ np.array([[1, .3, -.2], [.1, -1, .4], [.5, .2, .8]]))  # noqa
    X = cad_oracle.linear_estimation_rot(V, F, field(V), 1.0)
    # coincident vertices share one solution
    for c in np.unique(cls):
        m = np.nonzero(cls == c)[0]
        assert (X[m] == X[m[0]]).all()
    rng = np.random.default_rng(9)
    first = np.zeros(len(V), bool)
    first[np.unique(cls, return_index=True)[1]] = True
    u = rng.standard_normal(V.shape)
    u /= np.linalg.norm(u, axis=1, keepdims=True)
    errs = []
    for eps in (1e-5, 1e-6):
        Ve = V + eps * u * (~first)[:, None]          # copies moved, first members fixed
        assert cad_oracle.coincident_classes(Ve, F).max() + 1 == len(V)
        Xe = cad_oracle.linear_estimation_rot(Ve, F, field(Ve), 1.0)
        errs.append(np.abs(Xe - X).max())
    assert errs[1] < 0.2 * errs[0] and errs[1] < 1e-5
    # the refined solve satisfies the cancellation-free residual
    pairs = cad_oracle._rot_pairs(V, F, 1.0)
    A, B = cad_oracle.linear_system_rot(V, F, field(V), 1.0)
    nc = int(cls.max()) + 1
    PB = np.zeros((nc, 3))
    np.add.at(PB, cls, B)
    Xc = np.zeros((nc, 3))
    Xc[cls] = X
    res = PB - cad_oracle.linear_apply_rot(pairs, cls, Xc)
    # relative to each row's diagonal: pairs 5.6e-17 apart have weights of
    # 3.6e14, so one ulp of x leaves a residual of ~0.04 in their rows
    diag = np.bincount(cls, minlength=nc).astype(float)
    np.add.at(diag, cls[pairs[0]], pairs[2])
    np.add.at(diag, cls[pairs[1]], pairs[2])
    assert (np.abs(res).max(axis=1) / diag).max() < 1e-15


@pytest.mark.parametrize("grid", [(0.0, 1.0), (1.0, 0.0), (0.25, 0.8), (0.9, 0.1)])
def test_fused_step_times_match_the_torch_step(grid):
    """The (tau, sign, dt) handed to the fused HIP step are the network time
    inputs the float32 torch step feeds ODEFunc, and the step's sign and
    size: recorded from the torch path on the CPU."""
    torch.manual_seed(2)
    func = nf.ODEFunc()
    seen = []

    class Rec(torch.nn.Module):
        def forward(self, t, y):
            seen.append(float(t))
            return func(t, y)

    y = torch.rand(5, 3)
    t = torch.tensor(grid, dtype=torch.float32)
    y1 = nf.odeint(Rec(), y, t)[1]
    tau, sign, dt = nf.step_times(*grid)
    assert list(tau) == seen
    assert sign == (1.0 if grid[1] > grid[0] else -1.0)
    assert dt == pytest.approx(abs(grid[1] - grid[0]), rel=1e-6)
    # the same step written out with the recorded stage values
    k = []
    for s, tt in enumerate(tau):
        third = np.float32(1 / 3)
        z = (y if s == 0 else y + dt * k[0] * third if s == 1 else y + dt * (k[1] - k[0] * third) if s == 2
             else y + dt * (k[0] - k[1] + k[2]))
        k.append(sign * func(torch.tensor(tt), z))
    assert torch.equal(y1, y + (k[0] + 3 * (k[1] + k[2]) + k[3]) * dt * 0.125)
