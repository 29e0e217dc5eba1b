"""CAD chain of cad_deform (cad_deform.cc:21-120) on the C ABI.

* ``prepare``: RemoveDegenerated + MergeDuplex + Subdivide + geometry
  neighbours + representative graph (host code in libmeshode, as in the
  reference, mesh_tensor.cc:102-115).
* ``delaunay2`` / ``delaunay3``: the triangulations standing in for CGAL's
  (delaunay.cc:50-101).
* ``linear_estimation_rot``: LinearEstimationWithRot (linear.cc:114-230),
  frames, right-hand side and CG on the GPU.
* ``linear_estimation``: LinearEstimation (linear.cc:11-112), solved on the
  GPU.
* ``cad_deform``: the whole non-flow cad_deform pipeline on the GPU.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib


def _p(a):
    return C.c_void_p(a.ctypes.data) if a is not None and a.size else None


def prepare(V, F, len_subdivide=2e-2, step_neighbors=1.5e-2, step_graph=1e-2, threads=0):
    """dict(V, F, E, ref, GV, GE, clean_V, clean_F) of numpy arrays (FP64
    vertices, int32 indices); E/GE are (i < j) pairs in sorted order."""
    V = np.ascontiguousarray(V, np.float64).reshape(-1, 3)
    F = np.ascontiguousarray(F, np.int32).reshape(-1, 3)
    h = C.c_int64()
    _lib.call("mo_cad_prepare", _p(V), len(V), _p(F), len(F), len_subdivide, step_neighbors, step_graph,
              int(threads), C.byref(h))
    try:
        n = np.zeros(7, np.int64)
        _lib.call("mo_cad_sizes", h.value, _p(n))
        out = dict(V=np.zeros((n[0], 3)), F=np.zeros((n[1], 3), np.int32), E=np.zeros((n[2], 2), np.int32),
                   ref=np.zeros(n[0], np.int32), GV=np.zeros((n[3], 3)), GE=np.zeros((n[4], 2), np.int32),
                   clean_V=np.zeros((n[5], 3)), clean_F=np.zeros((n[6], 3), np.int32))
        _lib.call("mo_cad_export", h.value, *[_p(out[k]) for k in ("V", "F", "E", "ref", "GV", "GE", "clean_V",
                                                                   "clean_F")])
    finally:
        _lib.call("mo_cad_destroy", h.value)
    return out


def delaunay2(xy):
    xy = np.ascontiguousarray(xy, np.float64).reshape(-1, 2)
    cap = 2 * len(xy) + 8
    out = np.zeros((cap, 3), np.int32)
    n = C.c_int64()
    _lib.call("mo_delaunay2", _p(xy), len(xy), _p(out), cap, C.byref(n))
    return out[: n.value]


def delaunay3(xyz):
    """(dim, simplices [k, dim+1])."""
    xyz = np.ascontiguousarray(xyz, np.float64).reshape(-1, 3)
    cap = 8 * len(xyz) * len(xyz) + 16
    out = np.zeros(cap * 4, np.int32)
    n, dim = C.c_int64(), C.c_int32()
    _lib.call("mo_delaunay3", _p(xyz), len(xyz), _p(out), cap, C.byref(n), C.byref(dim))
    d = dim.value
    if d < 1:
        return d, np.zeros((0, max(d + 1, 1)), np.int32)
    return d, out[: n.value * (d + 1)].reshape(-1, d + 1)


def linear_estimation(V, F, E, ref, GV, rigidity=2.0):
    """LinearEstimation on the GPU.  V [n,3] FP64 CUDA tensor updated in place;
    F, E, ref int32 CUDA tensors; GV [g,3] FP64 CUDA tensor.  Returns the CG
    iteration count."""
    import torch
    for name, t in (("V", V), ("F", F), ("E", E), ("ref", ref), ("GV", GV)):
        if not t.is_cuda or not t.is_contiguous():
            raise ValueError(f"linear_estimation: {name} must be a contiguous device tensor")
    it = C.c_int32()
    _lib.call("mo_linear_estimation", C.c_void_p(V.data_ptr()), V.shape[0], C.c_void_p(F.data_ptr()), F.shape[0],
              C.c_void_p(E.data_ptr()), E.shape[0], C.c_void_p(ref.data_ptr()), C.c_void_p(GV.data_ptr()),
              GV.shape[0], float(rigidity), C.byref(it), C.c_void_p(torch.cuda.current_stream().cuda_stream))
    return it.value


def linear_estimation_rot(V, F, TV, rigidity=1.0):
    """LinearEstimationWithRot on the GPU.  V [n,3] FP64 CUDA tensor (rest
    positions) updated in place with the solution; F [m,3] int32 and TV [n,3]
    FP64 CUDA tensors.  Returns the CG iteration count."""
    import torch
    for name, t, dt, w in (("V", V, torch.float64, 3), ("F", F, torch.int32, 3), ("TV", TV, torch.float64, 3)):
        if not t.is_cuda or not t.is_contiguous() or t.dtype != dt or t.dim() != 2 or t.shape[1] != w:
            raise ValueError(f"linear_estimation_rot: {name} must be a contiguous [n,{w}] {dt} device tensor")
    if TV.shape[0] != V.shape[0]:
        raise ValueError("linear_estimation_rot: TV needs one row per vertex")
    it = C.c_int32()
    _lib.call("mo_linear_estimation_rot", C.c_void_p(V.data_ptr()), V.shape[0], C.c_void_p(F.data_ptr()),
              F.shape[0], C.c_void_p(TV.data_ptr()), float(rigidity), C.byref(it),
              C.c_void_p(torch.cuda.current_stream().cuda_stream))
    return it.value


_hip = None


def _copy_to_host(ptr, n):
    """[n,3] FP64 device array at ptr -> numpy (hipMemcpy, device to host)."""
    global _hip
    if _hip is None:
        _hip = C.CDLL("libamdhip64.so")
        _hip.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        _hip.hipMemcpy.restype = C.c_int
    out = np.empty((n, 3))
    rc = _hip.hipMemcpy(out.ctypes.data_as(C.c_void_p), C.c_void_p(ptr), out.nbytes, 2)
    if rc != 0:
        raise RuntimeError(f"hipMemcpy failed ({rc})")
    return out


def deform_subdivision(pid, V_rest, V, F, E, lam=1.0, options=None, max_solves=100000):
    """Deformer::DeformSubdivision with the flow callback (deformer.cc:
    173-263, cad_deform.cc:87-114): V (FP64 device tensor, normalized frame)
    is updated in place; returns (stats, snapshots) where snapshots are the
    positions handed to the callback (numpy, normalized frame); cad_deform
    appends the final state after the loop."""
    import torch
    from . import deform
    snaps = []

    def cb(ptr, n, user):
        snaps.append(_copy_to_host(ptr, n))

    cfn = _lib.FlowCallback(cb)
    st = _lib.LMStats()
    nsol = C.c_int32()
    opt = options or deform.default_options()
    stream = C.c_void_p(torch.cuda.current_stream().cuda_stream)
    _lib.call("mo_deform_subdivision", int(pid), C.c_void_p(V_rest.data_ptr()), C.c_void_p(V.data_ptr()),
              V.shape[0], C.c_void_p(F.data_ptr()), F.shape[0], C.c_void_p(E.data_ptr()), E.shape[0], float(lam),
              C.byref(opt), int(max_solves), cfn, None, C.byref(st), C.byref(nsol), stream)
    d = st.as_dict()
    d["solves"] = nsol.value
    return d, snaps


def cad_deform(cad_V, cad_F, ref_V, ref_F, grid_resolution=64, lam=1.0, symmetry=0, options=None):
    """The non-flow cad_deform pipeline (cad_deform.cc:30-120) in-process:
    host preprocessing, then on the GPU the reference's grid, DeformGraph on
    the representative graph and LinearEstimation.  Returns dict(V = deformed
    subdivided vertices in the reference's original frame (FP64 numpy),
    F, stats (DeformGraph summary), GV = deformed graph vertices (normalized
    frame), cg_iterations, prep = the preprocessing arrays)."""
    import torch
    from . import deform, objio  # noqa: F401
    ref_V = np.asarray(ref_V, np.float64).reshape(-1, 3)
    ref_F = np.asarray(ref_F, np.int32).reshape(-1, 3)
    if symmetry:
        n = len(ref_V)
        ref_V = np.concatenate([ref_V, ref_V * np.array([-1.0, 1.0, 1.0])])
        ref_F = np.concatenate([ref_F, ref_F + n]).astype(np.int32)
    prep = prepare(cad_V, cad_F)
    dev = torch.device("cuda", torch.cuda.current_device())
    pid = deform.template_f64(ref_V, ref_F, grid_resolution)
    try:
        stream = C.c_void_p(torch.cuda.current_stream().cuda_stream)
        V = torch.from_numpy(prep["V"]).to(dev).contiguous()
        GV = torch.from_numpy(prep["GV"]).to(dev).contiguous()
        F = torch.from_numpy(prep["F"]).to(dev).contiguous()
        E = torch.from_numpy(prep["E"]).to(dev).contiguous()
        R = torch.from_numpy(prep["ref"]).to(dev).contiguous()
        GE = torch.from_numpy(prep["GE"]).to(dev).contiguous()
        _lib.call("mo_apply_transform_f64", C.c_void_p(V.data_ptr()), V.shape[0], pid, 0, stream)
        _lib.call("mo_apply_transform_f64", C.c_void_p(GV.data_ptr()), GV.shape[0], pid, 0, stream)
        GV0 = GV.clone()
        stats = deform.deform_batch([pid], [GV], [GE], [1], [lam], options=options)[0]
        it = linear_estimation(V, F, E, R, GV, 2.0)
        _lib.call("mo_apply_transform_f64", C.c_void_p(V.data_ptr()), V.shape[0], pid, 1, stream)
        torch.cuda.synchronize()
        return dict(V=V.cpu().numpy(), F=prep["F"], stats=stats, GV=GV.cpu().numpy(), GV0=GV0.cpu().numpy(),
                    cg_iterations=it, prep=prep, pid=pid)
    except BaseException:
        _lib.call("mo_template_destroy", pid)
        raise
