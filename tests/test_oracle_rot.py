"""Pins the oracle's restatement of Deformer::DeformWithRot (deformer.cc:94-171,
EdgeLossWithRot edgeloss.h:64-98) on the CPU.

* The gradient J^T f (automatic differentiation through the angle-axis
  rotation, both of its branches) equals central differences of the cost.
* At the start (rest positions, zero rotations) every edge residual is 0:
  the cost is 1/2 sum d^2.
* With lambda = 0 the rotations see no force and the positions follow the
  plain Deformer::Deform solve exactly (same trust-region path).
* A target that is a rotated copy of the source is matched with rotations
  w = -theta axis (the rotation maps the deformed edge back onto the rest
  edge), which fixes the sign convention of the rotation formula.
"""
import numpy as np
import pytest

from meshode_amd import synth


def _problem(oracle, N=16, seed=9):
    _, _, tV, tF = synth.displaced_sphere_pair(seed, src_level=1, tgt_level=2)
    sV, sF = synth.icosphere(1)
    tn, sc, pos = oracle.normalize(tV)
    grid, _, _ = oracle.build_grid(tn, tF, N)
    sn = oracle.apply_transform(sV * 0.8, sc, pos)
    return grid, N, sn, sF


def test_rot_gradient_central_differences(oracle):
    grid, N, x0, F = _problem(oracle)
    rng = np.random.default_rng(1)
    x = x0 + rng.normal(0, 0.01, x0.shape)
    w = rng.normal(0, 0.2, x0.shape)
    w[0] = 0.0                       # first-order branch of the rotation
    w[1] = [3e-9, 0.0, 0.0]          # |w|^2 below double epsilon
    lam = 1.3
    g = oracle.lm_rot_gradient(grid, N, x, w, x0, F, lam)

    def cost(xx, ww):
        return oracle.lm_rot_cost(grid, N, xx, ww, x0, F, lam)[0]

    h = 1e-7
    for i in [0, 1] + list(rng.choice(len(x0), 6, replace=False)):
        for k in range(6):
            xp, xm, wp, wm = x.copy(), x.copy(), w.copy(), w.copy()
            if k < 3:
                xp[i, k] += h
                xm[i, k] -= h
            else:
                wp[i, k - 3] += h
                wm[i, k - 3] -= h
            fd = (cost(xp, wp) - cost(xm, wm)) / (2 * h)
            assert g[i, k] == pytest.approx(fd, rel=1e-4, abs=1e-6), (i, k)


def test_rot_start_cost_is_distance_only(oracle):
    grid, N, x0, F = _problem(oracle)
    tot, vc, ec = oracle.lm_rot_cost(grid, N, x0, np.zeros_like(x0), x0, F, 1.0)
    d = oracle.lookup_f64(grid, N, x0, want_grad=False)
    assert ec == 0.0
    assert tot == pytest.approx(0.5 * d @ d, rel=1e-14)
    _, _, st = oracle.lm_rot_solve(grid, N, x0, F, 1.0)
    assert st["initial_cost"] == pytest.approx(0.5 * d @ d, rel=1e-14)
    assert st["final_cost"] < st["initial_cost"]


def test_rot_lambda_zero_matches_plain_solve(oracle):
    grid, N, x0, F = _problem(oracle)
    x, w, st = oracle.lm_rot_solve(grid, N, x0, F, 0.0)
    xp, sp, _ = oracle.lm_solve(grid, N, x0, oracle.face_edges(F), 0.0)
    assert np.all(w == 0.0)
    np.testing.assert_allclose(x, xp, rtol=0, atol=1e-12)
    assert st["iterations"] == sp["iterations"]
    assert st["final_cost"] == pytest.approx(sp["final_cost"], rel=1e-12)


def test_rot_recovers_rotated_target(oracle):
    """Target = the source rotated by 12 degrees about an axis through its
    centre: the solve reaches the rotated pose (within a grid cell) and the
    rotations point along -theta axis."""
    _, _, tV, tF = synth.displaced_sphere_pair(3, src_level=2, tgt_level=2)
    theta = np.deg2rad(12.0)
    axis = np.array([0.3, -0.5, 0.8])
    axis /= np.linalg.norm(axis)
    K = np.array([[0, -axis[2], axis[1]], [axis[2], 0, -axis[0]], [-axis[1], axis[0], 0]])
    R = np.eye(3) + np.sin(theta) * K + (1 - np.cos(theta)) * K @ K
    c = tV.mean(0)
    rV = (tV - c) @ R.T + c
    N = 32
    tn, sc, pos = oracle.normalize(rV)
    grid, _, _ = oracle.build_grid(tn, tF, N)
    src = oracle.apply_transform(tV, sc, pos)
    x, w, st = oracle.lm_rot_solve(grid, N, src, tF, 1.0)
    # the exact rotated pose has zero rigidity cost; its distance cost is the
    # trilinear field's floor near the surface, which the solve reaches
    cn = src.mean(0)
    ideal = (src - cn) @ R.T + cn
    tot, _, ec = oracle.lm_rot_cost(grid, N, ideal, np.tile(-theta * axis, (len(src), 1)), src, tF, 1.0)
    assert ec < 1e-25
    assert st["final_cost"] < 1.01 * tot and st["final_cost"] < 0.5 * st["initial_cost"]
    assert np.abs(x - ideal).max() < 1.0 / N
    mean_w = w.mean(0)
    assert mean_w @ axis < -0.5 * theta
    assert np.linalg.norm(mean_w - (mean_w @ axis) * axis) < 0.35 * theta
