"""ctypes wrapper for the CPU oracle (oracle/liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, by ``__graft_entry__.smoke()``
as the checker, and by ``bench.py``'s cpu_baseline leg.  The product path
(meshode_amd) never imports this module.

Each wrapper names the reference function it restates; the arithmetic lives
in meshode_oracle.cpp next to the citations.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle.so")
_lib = None


class LMOptions(C.Structure):
    _fields_ = [
        ("max_num_iterations", C.c_int),
        ("max_consecutive_invalid_steps", C.c_int),
        ("function_tolerance", C.c_double),
        ("gradient_tolerance", C.c_double),
        ("parameter_tolerance", C.c_double),
        ("min_relative_decrease", C.c_double),
        ("initial_trust_region_radius", C.c_double),
        ("max_trust_region_radius", C.c_double),
        ("min_trust_region_radius", C.c_double),
        ("min_lm_diagonal", C.c_double),
        ("max_lm_diagonal", C.c_double),
    ]


class LMStats(C.Structure):
    _fields_ = [
        ("initial_cost", C.c_double),
        ("final_cost", C.c_double),
        ("vertex_cost", C.c_double),
        ("rigidity_cost", C.c_double),
        ("iterations", C.c_int),
        ("successful_steps", C.c_int),
        ("unsuccessful_steps", C.c_int),
        ("termination", C.c_int),
    ]


def build() -> str:
    """Compile liboracle.so with the committed Makefile (g++, no GPU)."""
    subprocess.run(["make", "-s", "-C", _HERE], check=True)
    return _LIB_PATH


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        _lib = C.CDLL(_LIB_PATH)
        _lib.mo_oracle_lm_solve.restype = C.c_int
        _lib.mo_oracle_lm_cost.restype = C.c_double
        _lib.mo_oracle_lm_cost_ex.restype = C.c_double
        _lib.mo_oracle_lm_solve_ex.restype = C.c_int
        _lib.mo_oracle_lm_rot_solve.restype = C.c_int
        _lib.mo_oracle_lm_rot_cost.restype = C.c_double
    return _lib


def _p(a):
    return a.ctypes.data_as(C.c_void_p) if a is not None else None


def _f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def _f32(a):
    return np.ascontiguousarray(a, dtype=np.float32)


def _i32(a):
    return np.ascontiguousarray(a, dtype=np.int32)


def normalize(V):
    """Mesh::Normalize (mesh.cc:66-96). Returns (normalized V, scale, pos)."""
    V = _f64(V).copy()
    scale = C.c_double()
    pos = np.zeros(3)
    lib().mo_oracle_normalize(_p(V), C.c_long(len(V)), C.byref(scale), _p(pos))
    return V, scale.value, pos


def apply_transform(V, scale, pos):
    """Mesh::ApplyTransform (mesh.cc:98-105)."""
    V = _f64(V).copy()
    lib().mo_oracle_apply_transform(_p(V), C.c_long(len(V)), C.c_double(scale), _p(_f64(pos)))
    return V


def normalize_by_template_f32(V, scale, trans):
    """NormalizeByTemplate (normalize.cc:5-24) on float32."""
    V = _f32(V).copy()
    lib().mo_oracle_normalize_by_template_f32(_p(V), C.c_long(len(V)), C.c_double(scale), _p(_f64(trans)))
    return V


def denormalize_by_template_f32(V, scale, trans):
    """DenormalizeByTemplate (normalize.cc:26-45) on float32."""
    V = _f32(V).copy()
    lib().mo_oracle_denormalize_by_template_f32(_p(V), C.c_long(len(V)), C.c_double(scale), _p(_f64(trans)))
    return V


def reflection_symmetrize(V, F):
    """Mesh::ReflectionSymmetrize (mesh.cc:234-246)."""
    V = _f64(V); F = _i32(F)
    n, m = len(V), len(F)
    V2 = np.zeros((2 * n, 3)); V2[:n] = V
    F2 = np.zeros((2 * m, 3), np.int32); F2[:m] = F
    lib().mo_oracle_reflection_symmetrize(_p(V2), C.c_long(n), _p(F2), C.c_long(m))
    return V2, F2


def point_mesh_d2(Q, V, F):
    """Squared distance and nearest face (lowest index on ties)."""
    Q = _f64(Q); V = _f64(V); F = _i32(F)
    d2 = np.zeros(len(Q)); idx = np.zeros(len(Q), np.int32)
    lib().mo_oracle_point_mesh_d2(_p(Q), C.c_long(len(Q)), _p(V), _p(F), C.c_long(len(F)), _p(d2), _p(idx))
    return d2, idx


def build_grid(V, F, N, cells=None):
    """Mesh::ConstructDistanceField (mesh.cc:106-152).

    Returns (dist f64, dist f32, nearest i32), flat over (z, y, x) or over
    ``cells`` when given."""
    V = _f64(V); F = _i32(F)
    if cells is None:
        n = N * N * N
        cp = None
    else:
        cells = np.ascontiguousarray(cells, dtype=np.int64)
        n = len(cells)
        cp = _p(cells)
    d = np.zeros(n); d32 = np.zeros(n, np.float32); idx = np.zeros(n, np.int32)
    lib().mo_oracle_build_grid(_p(V), _p(F), C.c_long(len(F)), C.c_int(N), cp, C.c_long(n),
                               _p(d), _p(d32), _p(idx))
    return d, d32, idx


def lookup_band2(N):
    """Squared band of a template grid: 0.2 + sqrt(3)/N + 0.01, squared (the
    lookups truncate interpolated distances above 0.2, uniformgrid.cc:80-82,
    and the field is 1-Lipschitz, so values beyond the band never reach a
    lookup result)."""
    b = 0.2 + 1.7320508075688772 / N + 0.01
    return b * b


def build_grid_banded(V, F, N, cells=None):
    """The template grid: per grid point sqrt(min(d^2, band^2)) and the
    nearest face (-1 when none is within the band).  Returns (dist64, dist32,
    nearest) over (z, y, x) or over ``cells``."""
    if cells is None:
        cells = np.arange(N ** 3)
    cells = np.asarray(cells, np.int64)
    i, rem = np.divmod(cells, N * N)
    j, k = np.divmod(rem, N)
    Q = np.stack([k / N, j / N, i / N], 1).astype(np.float64)
    d2, idx = point_mesh_d2(Q, V, F)
    b2 = lookup_band2(N)
    inside = d2 <= b2
    d = np.where(inside, np.sqrt(np.where(inside, d2, b2)), np.sqrt(b2))
    return d, d.astype(np.float32), np.where(inside, idx, -1).astype(np.int32)


def lookup_f32(grid, N, P, want_grad=True):
    """UniformGrid::DistanceFloat<float / Jet<float,3>> (uniformgrid.cc:85-150)."""
    grid = _f64(grid); P = _f32(P)
    d = np.zeros(len(P), np.float32)
    g = np.zeros((len(P), 3), np.float32) if want_grad else None
    lib().mo_oracle_lookup_f32(_p(grid), C.c_int(N), _p(P), C.c_long(len(P)), _p(d), _p(g))
    return (d, g) if want_grad else d


def lookup_f64(grid, N, P, want_grad=True):
    """UniformGrid::distance<double / Jet<double,3>> (uniformgrid.cc:18-83)."""
    grid = _f64(grid); P = _f64(P)
    d = np.zeros(len(P))
    g = np.zeros((len(P), 3)) if want_grad else None
    lib().mo_oracle_lookup_f64(_p(grid), C.c_int(N), _p(P), C.c_long(len(P)), _p(d), _p(g))
    return (d, g) if want_grad else d


def distance_loss_fwd(grid, N, V):
    """DistanceFieldLoss_forward (distance_layer.cc:8-37)."""
    grid = _f64(grid); V = _f32(V)
    out = np.zeros(len(V), np.float32)
    lib().mo_oracle_distance_loss_fwd(_p(grid), C.c_int(N), _p(V), C.c_long(len(V)), _p(out))
    return out


def distance_loss_bwd(grid, N, V):
    """DistanceFieldLoss_backward (distance_layer.cc:39-80)."""
    grid = _f64(grid); V = _f32(V)
    out = np.zeros((len(V), 3), np.float32)
    lib().mo_oracle_distance_loss_bwd(_p(grid), C.c_int(N), _p(V), C.c_long(len(V)), _p(out))
    return out


def store_rigidity(V, F):
    """StoreRigidityInformation (rigid_layer.cc:11-46)."""
    V = _f32(V); F = _i32(F)
    off = np.zeros((3 * len(F), 3), np.float32)
    lib().mo_oracle_store_rigidity(_p(V), _p(F), C.c_long(len(F)), _p(off))
    return off


def rigid_fwd(V, F, off):
    """RigidEdgeLoss_forward (rigid_layer.cc:48-89)."""
    V = _f32(V); F = _i32(F); off = _f32(off)
    out = np.zeros((3 * len(F), 3), np.float32)
    lib().mo_oracle_rigid_fwd(_p(V), _p(F), C.c_long(len(F)), _p(off), _p(out))
    return out


def rigid_bwd(V, F, off):
    """RigidEdgeLoss_backward (rigid_layer.cc:91-133)."""
    V = _f32(V); F = _i32(F); off = _f32(off)
    out = np.zeros((len(V), 3), np.float32)
    lib().mo_oracle_rigid_bwd(_p(V), C.c_long(len(V)), _p(F), C.c_long(len(F)), _p(off), _p(out))
    return out


def store_graph(V, E):
    """StoreGraphInformation (graph_layer.cc:11-45)."""
    V = _f32(V); E = _i32(E)
    off = np.zeros((len(E), 3), np.float32)
    lib().mo_oracle_store_graph(_p(V), _p(E), C.c_long(len(E)), _p(off))
    return off


def graph_fwd(V, E, off):
    """GraphEdgeLoss_forward (graph_layer.cc:47-91)."""
    V = _f32(V); E = _i32(E); off = _f32(off)
    out = np.zeros((len(E), 3), np.float32)
    lib().mo_oracle_graph_fwd(_p(V), _p(E), C.c_long(len(E)), _p(off), _p(out))
    return out


def graph_bwd(V, E, off):
    """GraphEdgeLoss_backward (graph_layer.cc:93-135)."""
    V = _f32(V); E = _i32(E); off = _f32(off)
    out = np.zeros((len(V), 3), np.float32)
    lib().mo_oracle_graph_bwd(_p(V), C.c_long(len(V)), _p(E), C.c_long(len(E)), _p(off), _p(out))
    return out


def default_lm_options() -> LMOptions:
    o = LMOptions()
    lib().mo_oracle_lm_default_options(C.byref(o))
    return o


def face_edges(F):
    """Directed edges of Deformer::Deform (deformer.cc:40-53): (F[i][j], F[i][(j+1)%3])."""
    F = _i32(F)
    a = F.reshape(-1)
    b = F[:, [1, 2, 0]].reshape(-1)
    return np.stack([a, b], 1).astype(np.int32)


def lm_solve(grid, N, V, E, lam=1.0, options: LMOptions | None = None, trace=False):
    """Deformer::Deform / DeformGraph (deformer.cc:18-92, 370-442).

    V: [n,3] FP64 normalized positions (not modified); E: directed edges.
    Returns (V_out, stats dict, trace array or None)."""
    grid = _f64(grid); x = _f64(V).copy(); E = _i32(E)
    opt = options or default_lm_options()
    st = LMStats()
    tr = np.full((opt.max_num_iterations, 8), np.nan) if trace else None
    lib().mo_oracle_lm_solve(_p(grid), C.c_int(N), _p(x), C.c_int(len(x)), _p(E), C.c_int(len(E)),
                             C.c_double(lam), C.byref(opt), C.byref(st), _p(tr))
    stats = {f: getattr(st, f) for f, _ in LMStats._fields_}
    if tr is not None:
        tr = tr[: stats["iterations"]]
    return x, stats, tr


def lm_solve_ex(grid, N, x, x0, E, lam_e=None, lam=1.0, max_successful=0, options: LMOptions | None = None):
    """Deformer::DeformSubdivision-style solve (deformer.cc:173-263): start
    at x, rest offsets from x0, per-edge weights lam_e (AdaptiveEdgeLoss,
    edgeloss.h:35-61) or the constant lam, and at most max_successful
    successful steps (TerminateWhenSuccessCallback, callback.h:22-36).
    Returns (x_out, stats)."""
    grid = _f64(grid); xo = _f64(x).copy(); x0 = _f64(x0); E = _i32(E)
    le = None if lam_e is None else _f64(lam_e)
    opt = options or default_lm_options()
    st = LMStats()
    lib().mo_oracle_lm_solve_ex(_p(grid), C.c_int(N), _p(xo), _p(x0), C.c_int(len(xo)), _p(E), C.c_int(len(E)),
                                _p(le) if le is not None else None, C.c_double(lam), C.c_int(max_successful),
                                C.byref(opt), C.byref(st), None)
    return xo, {f: getattr(st, f) for f, _ in LMStats._fields_}


def lm_cost_ex(grid, N, x, x0, E, lam_e=None, lam=1.0):
    grid = _f64(grid); x = _f64(x); x0 = _f64(x0); E = _i32(E)
    le = None if lam_e is None else _f64(lam_e)
    vc = C.c_double(); ec = C.c_double()
    tot = lib().mo_oracle_lm_cost_ex(_p(grid), C.c_int(N), _p(x), _p(x0), C.c_int(len(x)), _p(E), C.c_int(len(E)),
                                     _p(le) if le is not None else None, C.c_double(lam), C.byref(vc), C.byref(ec))
    return tot, vc.value, ec.value


def lm_rot_solve(grid, N, V, F, lam=1.0, W=None, options: LMOptions | None = None):
    """Deformer::DeformWithRot (deformer.cc:94-171): positions V [n,3] FP64
    in the normalized frame (start and rest shape; not modified), rotations
    W [n,3] (zeros by default, as the reference), faces F.
    Returns (V_out, W_out, stats dict)."""
    grid = _f64(grid); x = _f64(V).copy(); F = _i32(F)
    w = np.zeros_like(x) if W is None else _f64(W).copy()
    opt = options or default_lm_options()
    st = LMStats()
    lib().mo_oracle_lm_rot_solve(_p(grid), C.c_int(N), _p(x), _p(w), C.c_int(len(x)), _p(F), C.c_int(len(F)),
                                 C.c_double(lam), C.byref(opt), C.byref(st))
    return x, w, {f: getattr(st, f) for f, _ in LMStats._fields_}


def lm_rot_cost(grid, N, x, w, x0, F, lam=1.0):
    """Vertex / rigidity / total cost of DeformWithRot (deformer.cc:152-170)
    at positions x and rotations w with rest offsets from x0."""
    grid = _f64(grid); x = _f64(x); w = _f64(w); x0 = _f64(x0); F = _i32(F)
    vc = C.c_double(); ec = C.c_double()
    tot = lib().mo_oracle_lm_rot_cost(_p(grid), C.c_int(N), _p(x), _p(w), _p(x0), C.c_int(len(x)), _p(F),
                                      C.c_int(len(F)), C.c_double(lam), C.byref(vc), C.byref(ec))
    return tot, vc.value, ec.value


def lm_rot_gradient(grid, N, x, w, x0, F, lam=1.0):
    """J^T f of DeformWithRot at (x, w), [n,6]: d/dp then d/dw."""
    grid = _f64(grid); x = _f64(x); w = _f64(w); x0 = _f64(x0); F = _i32(F)
    g = np.zeros((len(x), 6))
    lib().mo_oracle_lm_rot_gradient(_p(grid), C.c_int(N), _p(x), _p(w), _p(x0), C.c_int(len(x)), _p(F),
                                    C.c_int(len(F)), C.c_double(lam), _p(g))
    return g


def adaptive_lambda(V, E, lam=1.0):
    """AdaptiveEdgeLoss's weight (edgeloss.h:36-40): lambda * 2e-2 / (|v| + 1e-8),
    v = V[a] - V[b] (FP64)."""
    V = np.asarray(V, np.float64)
    E = np.asarray(E)
    v = V[E[:, 0]] - V[E[:, 1]]
    return lam * (2e-2 / (np.sqrt(v[:, 0] * v[:, 0] + v[:, 1] * v[:, 1] + v[:, 2] * v[:, 2]) + 1e-8))


def lm_cost(grid, N, x, x0, E, lam=1.0):
    """Vertex / rigidity / total cost of Deformer::Deform (deformer.cc:76-91)."""
    grid = _f64(grid); x = _f64(x); x0 = _f64(x0); E = _i32(E)
    vc = C.c_double(); ec = C.c_double()
    tot = lib().mo_oracle_lm_cost(_p(grid), C.c_int(N), _p(x), _p(x0), C.c_int(len(x)), _p(E),
                                  C.c_int(len(E)), C.c_double(lam), C.byref(vc), C.byref(ec))
    return tot, vc.value, ec.value


def build_grid_bvh(V, F, N):
    """Same as build_grid (identical outputs) with an AABB tree over the faces."""
    V = _f64(V); F = _i32(F)
    n = N * N * N
    d = np.zeros(n); d32 = np.zeros(n, np.float32); idx = np.zeros(n, np.int32)
    lib().mo_oracle_build_grid_bvh(_p(V), _p(F), C.c_long(len(F)), C.c_int(N), _p(d), _p(d32), _p(idx))
    return d, d32, idx
