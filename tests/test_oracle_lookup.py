"""Pins the oracle's trilinear lookups (UniformGrid::distance<T> and
DistanceFloat<T>, uniformgrid.cc:18-150), their Jet gradients and the
distance loss layer (distance_layer.cc:8-80) on the CPU: affine fields are
reproduced exactly by trilinear interpolation, the 0.2 truncation, the
out-of-box penalties (including the index N-1 zero-penalty quirk),
truncation of the cell index toward zero, and central differences."""
import numpy as np
import pytest

N = 16


def _affine_grid(a=(0.05, 0.03, 0.02), b=0.01, n=N):
    k = np.arange(n) / n
    Z, Y, X = np.meshgrid(k, k, k, indexing="ij")
    return (b + a[0] * X + a[1] * Y + a[2] * Z).reshape(-1)


def test_affine_field_reproduced(oracle):
    g = _affine_grid()
    rng = np.random.default_rng(0)
    P = rng.uniform(0.0, (N - 1) / N - 1e-6, (2000, 3))
    f = 0.01 + P @ np.array([0.05, 0.03, 0.02])
    d64, g64 = oracle.lookup_f64(g, N, P)
    assert np.allclose(d64, f, rtol=1e-13, atol=1e-15)
    assert np.allclose(g64, [0.05, 0.03, 0.02], rtol=1e-10, atol=1e-12)
    d32, g32 = oracle.lookup_f32(g, N, P.astype(np.float32))
    assert np.allclose(d32, f, rtol=2e-6, atol=1e-7)
    assert np.allclose(g32, [0.05, 0.03, 0.02], rtol=2e-4, atol=1e-5)


def test_truncation_above_0_2(oracle):
    g = np.full(N ** 3, 0.3)
    P = np.random.default_rng(1).uniform(0.1, 0.8, (100, 3))
    d, gr = oracle.lookup_f64(g, N, P)
    assert np.all(d == 0) and np.all(gr == 0)
    d32, g32 = oracle.lookup_f32(g, N, P.astype(np.float32))
    assert np.all(d32 == 0) and np.all(g32 == 0)


def test_out_of_box_penalties(oracle):
    g = _affine_grid()
    P = np.array([
        [-0.1, 0.5, 0.5],              # px < 0: -p N
        [1.2, 0.5, 0.5],               # px >= N: p N - (N - 1 - 1e-3)
        [(N - 0.5) / N, 0.5, 0.5],     # px = N-1: out-of-box branch, zero penalty
        [-0.1, 1.2, 0.5],              # two axes
    ])
    d, gr = oracle.lookup_f64(g, N, P)
    assert d[0] == pytest.approx(0.1 * N)
    assert d[1] == pytest.approx(1.2 * N - (N - 1 - 1e-3))
    assert d[2] == 0.0
    assert d[3] == pytest.approx(0.1 * N + 1.2 * N - (N - 1 - 1e-3))
    assert np.array_equal(gr[0], [-N, 0, 0])
    assert np.array_equal(gr[1], [N, 0, 0])
    assert np.array_equal(gr[2], [0, 0, 0])
    assert np.array_equal(gr[3], [-N, N, 0])
    d32, g32 = oracle.lookup_f32(g, N, P.astype(np.float32))
    assert d32[0] == pytest.approx(0.1 * N, rel=1e-6) and d32[2] == 0.0
    assert np.array_equal(g32[3], [-N, N, 0])


def test_cell_index_truncates_toward_zero(oracle):
    # p*N in (-1, 0) truncates to 0: interior branch, linear extrapolation
    g = _affine_grid()
    P = np.array([[-0.3 / N, 0.5, 0.5]])
    d, _ = oracle.lookup_f64(g, N, P)
    assert d[0] == pytest.approx(0.01 + 0.05 * (-0.3 / N) + 0.03 * 0.5 + 0.02 * 0.5, rel=1e-12)


def test_jet_gradient_central_differences(oracle):
    rng = np.random.default_rng(2)
    g = rng.uniform(0.0, 0.15, N ** 3)
    P = rng.uniform(0.05, 0.9, (3000, 3))
    frac = P * N - np.floor(P * N)
    keep = np.all((frac > 1e-3) & (frac < 1 - 1e-3), axis=1)
    P = P[keep]
    d, gr = oracle.lookup_f64(g, N, P)
    keep = (d > 1e-6) & (d < 0.2 - 1e-4)
    P, d, gr = P[keep], d[keep], gr[keep]
    h = 1e-7
    fd = np.zeros_like(P)
    for k in range(3):
        e = np.zeros(3); e[k] = h
        fd[:, k] = (oracle.lookup_f64(g, N, P + e, want_grad=False) - oracle.lookup_f64(g, N, P - e, want_grad=False)) / (2 * h)
    assert np.allclose(gr, fd, rtol=1e-6, atol=1e-8)


def test_f32_matches_f64_semantics(oracle):
    rng = np.random.default_rng(3)
    g = rng.uniform(0.0, 0.19, N ** 3)
    P = rng.uniform(0.0, 0.93, (4000, 3)).astype(np.float32)
    d32 = oracle.lookup_f32(g, N, P, want_grad=False)
    d64 = oracle.lookup_f64(g, N, P.astype(np.float64), want_grad=False)
    # the cell index may differ where p*N is within float rounding of an
    # integer (uniformgrid.cc:20-22 vs 87-89); elsewhere the values agree
    frac = P.astype(np.float64) * N - np.floor(P.astype(np.float64) * N)
    ok = np.all((frac > 1e-5) & (frac < 1 - 1e-5), axis=1)
    assert np.allclose(d32[ok], d64[ok], rtol=1e-5, atol=1e-6)


def test_distance_loss_layer(oracle):
    rng = np.random.default_rng(4)
    g = rng.uniform(0.0, 0.15, N ** 3)
    V = rng.uniform(-0.1, 1.1, (2000, 3)).astype(np.float32)
    d, gr = oracle.lookup_f32(g, N, V)
    fwd = oracle.distance_loss_fwd(g, N, V)
    bwd = oracle.distance_loss_bwd(g, N, V)
    assert np.array_equal(fwd, d * d)
    # 0.5 * d(d^2)/dp = d * grad d, rounded once to float
    assert np.array_equal(bwd, (d[:, None] * gr).astype(np.float32))


def test_banded_grid_gives_identical_lookups(oracle):
    """Template grids hold sqrt(min(d^2, band^2)), band = 0.2 + sqrt(3)/N +
    0.01: every lookup (FP32 and FP64 paths, value and gradient) equals the
    lookup on the exact field, on probes spread over the box and
    concentrated where the field crosses 0.2 and the band."""
    from meshode_amd import synth
    _, _, V, F = synth.displaced_sphere_pair(4, src_level=1, tgt_level=2)
    for N in (12, 24):
        Vn, _, _ = oracle.normalize(V)
        full, _, idx = oracle.build_grid(Vn, F, N)
        band, band32, bidx = oracle.build_grid_banded(Vn, F, N)
        b = np.sqrt(oracle.lookup_band2(N))
        inside = full < b
        assert np.array_equal(band[inside], full[inside]) and np.array_equal(bidx[inside], idx[inside])
        assert np.all(band[~inside] == b) and np.all(bidx[~inside] == -1) and (~inside).any()
        rng = np.random.default_rng(N)
        P = rng.uniform(-0.05, 1.05, (40000, 3))
        # probes where the interpolated field is near 0.2 or the band
        d = oracle.lookup_f64(full, N, P, want_grad=False)
        near = P[(np.abs(d - 0.2) < 0.03) | (d == 0)][:20000]
        P = np.concatenate([P, near + rng.normal(0, 1e-3, near.shape)])
        for fn in (oracle.lookup_f32, oracle.lookup_f64):
            a_d, a_g = fn(full, N, P)
            b_d, b_g = fn(band, N, P)
            assert np.array_equal(a_d, b_d) and np.array_equal(a_g, b_g)
        assert (oracle.lookup_f64(full, N, P, want_grad=False) == 0).sum() > 1000   # the truncation is exercised
