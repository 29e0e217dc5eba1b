"""Drop-in replacement for the reference's ``pyDeform`` module on MI355X.

Same function names, positional signatures and semantics as
``PYBIND11_MODULE(pyDeform, m)`` (src/interface/pydeform.cc:14-39); every
computation runs in libmeshode.so (HIP, gfx950) through its C ABI.

Differences from the reference, all on inputs the reference leaves
undefined:
  * tensors may live on the CPU or on a ROCm device; results come back on
    the input's device (the reference requires CPU tensors,
    distance_layer.cc:21);
  * non-contiguous tensors are handled through ``.contiguous()`` (the
    reference reads ``storage().data()`` and ignores offsets and strides);
  * wrong dtypes, unknown ``param_id`` values, out-of-range indices and
    loss calls before the matching Store call raise instead of reading
    out of bounds.
``LoadCadMesh`` runs the CAD preprocessing chain (host code in libmeshode,
as in the reference) and ``SolveLinear`` solves LinearEstimation
(with_rot = 0) or LinearEstimationWithRot (with_rot = 1) on the GPU.  The
``Cad*`` layer functions (cad_layer.cc, whose only Python caller is broken,
cad_loss_layer.py:10-11) raise ``NotImplementedError``.
"""
from __future__ import annotations

import ctypes as C

import torch

from . import _lib
from . import objio

__all__ = [
    "LoadMesh", "LoadCadMesh", "SaveMesh",
    "InitializeDeformTemplate", "NormalizeByTemplate", "DenormalizeByTemplate", "SolveLinear",
    "DistanceFieldLoss_forward", "DistanceFieldLoss_backward",
    "RigidEdgeLoss_forward", "RigidEdgeLoss_backward", "StoreRigidityInformation",
    "CadEdgeLoss_forward", "CadEdgeLoss_backward", "StoreCadInformation",
    "GraphEdgeLoss_forward", "GraphEdgeLoss_backward", "StoreGraphInformation",
]


def _device():
    if not torch.cuda.is_available():
        raise RuntimeError("pyDeform (MI355X) needs a ROCm GPU: torch.cuda.is_available() is False")
    return torch.device("cuda", torch.cuda.current_device())


def _stream():
    return C.c_void_p(torch.cuda.current_stream().cuda_stream)


def _ptr(t):
    return C.c_void_p(t.data_ptr()) if t is not None and t.numel() > 0 else C.c_void_p(0)


def _check(t, dtype, name, width):
    if not isinstance(t, torch.Tensor):
        raise TypeError(f"{name} must be a torch.Tensor")
    if t.dtype != dtype:
        raise TypeError(f"{name} must be {dtype} (got {t.dtype})")
    if t.dim() != 2 or t.shape[1] != width:
        if not (t.numel() == 0):
            raise ValueError(f"{name} must have shape [n,{width}] (got {tuple(t.shape)})")


def _to_dev(t):
    """Contiguous device view of t (a copy when t is on the CPU)."""
    dev = _device()
    if t.device != dev:
        return t.detach().to(dev).contiguous()
    return t.detach().contiguous()


# Device copies of CPU index tensors, so that an Adam loop calling the edge
# layers with the same F/E every step uploads it once.  An entry is keyed by
# storage pointer, size, shape and version counter and holds the source
# tensor itself: while the entry lives, that storage cannot be freed and its
# address reused by another tensor, and an in-place change bumps _version.
_index_cache: dict = {}


def _index_to_dev(t):
    dev = _device()
    if t.device == dev:
        return t.contiguous()
    key = (t.data_ptr(), t.numel(), tuple(t.shape), t._version, str(t.dtype))
    hit = _index_cache.get(key)
    if hit is not None:
        return hit[1]
    d = t.detach().to(dev).contiguous()
    if len(_index_cache) > 32:
        _index_cache.clear()
    _index_cache[key] = (t, d)
    return d


def _nrows(t):
    return int(t.shape[0]) if t.dim() > 0 else 0


# ---------------------------------------------------------------------------
# Mesh I/O (mesh_tensor.cc:87-100, 180-186; mesh.cc:14-64)

def LoadMesh(filename):
    """Returns [V float32 [n,3] (raw, not normalized), F int32 [m,3]]."""
    V, F = objio.read_obj(filename)
    return [torch.from_numpy(V.astype("float32")), torch.from_numpy(F.astype("int32"))]


def SaveMesh(filename, tensorV, tensorF):
    """Writes `v x y z` / 1-based `f a b c` lines (Mesh::WriteOBJ)."""
    V = tensorV.detach().cpu().contiguous().double().numpy()
    F = tensorF.detach().cpu().contiguous().numpy()
    objio.write_obj(filename, V, F)


def LoadCadMesh(filename):
    """mesh_tensor.cc:102-178: read the OBJ, RemoveDegenerated, MergeDuplex,
    Subdivide(2e-2), ComputeGeometryNeighbors(1.5e-2),
    ComputeRepresentativeGraph(1e-2) (host code in libmeshode).  Returns
    [V f32 [n,3] (raw coordinates), F i32 [m,3], E i32 [e,2] (geometry
    neighbour pairs), Src2Graph i32 [n,1], GraphV f32 [g,3], GraphE i32
    [h,2]] as CPU tensors, like the reference."""
    from . import cad
    V, F = objio.read_obj(filename)
    r = cad.prepare(V, F)
    return [torch.from_numpy(r["V"].astype("float32")), torch.from_numpy(r["F"]),
            torch.from_numpy(r["E"]), torch.from_numpy(r["ref"].reshape(-1, 1).copy()),
            torch.from_numpy(r["GV"].astype("float32")), torch.from_numpy(r["GE"])]


def SolveLinear(tensorV, tensorF, tensorE, tensorRef, tensorGraphV, rigidity, with_rot=0):
    """linear_layer.cc:5-87.  with_rot = 0: LinearEstimation (linear.cc:11-112)
    with E, the per-vertex graph references and the graph vertices.
    with_rot = 1: LinearEstimationWithRot (linear.cc:114-230); tensorGraphV
    holds one target position per vertex and E / references are unused, as
    in the reference.  Both are solved on the GPU in FP64; tensorV (float32
    [n,3]) is overwritten with the solution, stored as float32, on its own
    device.  The reference binding has no default for with_rot, so its
    6-argument call from graph_loss_layer.py's Finalize raises; here
    with_rot defaults to 0."""
    from . import cad
    dev = _device()
    V = tensorV.detach().to(dev).reshape(-1, 3).double().contiguous()
    F = tensorF.detach().to(dev, torch.int32).reshape(-1, 3).contiguous()
    GV = tensorGraphV.detach().to(dev).reshape(-1, 3).double().contiguous()
    if with_rot:
        if GV.shape[0] != V.shape[0]:
            raise ValueError("SolveLinear(with_rot=1): one target position per vertex is required")
        cad.linear_estimation_rot(V, F, GV, float(rigidity))
    else:
        E = tensorE.detach().to(dev, torch.int32).reshape(-1, 2).contiguous()
        R = tensorRef.detach().to(dev, torch.int32).reshape(-1).contiguous()
        if R.numel() != V.shape[0]:
            raise ValueError("SolveLinear: one graph reference per vertex is required")
        cad.linear_estimation(V, F, E, R, GV, float(rigidity))
    with torch.no_grad():
        tensorV.copy_(V.to(tensorV.device, torch.float32).reshape(tensorV.shape))


def CadEdgeLoss_forward(*args):
    raise NotImplementedError("CadEdgeLoss_* is outside the deformation hot path of this build")


CadEdgeLoss_backward = CadEdgeLoss_forward
StoreCadInformation = CadEdgeLoss_forward


# ---------------------------------------------------------------------------
# Templates (deform_params.cc:16-40, normalize.cc:5-45)

def InitializeDeformTemplate(tensorV, tensorF, symmetry, grid_resolution):
    """Normalizes the target and builds its grid_resolution^3 distance grid.
    Returns the integer param_id."""
    _check(tensorV, torch.float32, "tensorV", 3)
    _check(tensorF, torch.int32, "tensorF", 3)
    V = _to_dev(tensorV)
    F = _to_dev(tensorF)
    pid = C.c_int32(-1)
    _lib.call("mo_template_create_f32", _ptr(V), _nrows(V), _ptr(F), _nrows(F), int(symmetry),
              int(grid_resolution), _stream(), C.byref(pid))
    return int(pid.value)


def _inplace(name, tensorV, param_id):
    _check(tensorV, torch.float32, "tensorV", 3)
    V = _to_dev(tensorV)
    _lib.call(name, _ptr(V), _nrows(V), int(param_id), _stream())
    if V.data_ptr() != tensorV.data_ptr():
        tensorV.data.copy_(V.to(tensorV.device).view_as(tensorV))


def NormalizeByTemplate(tensorV, param_id):
    _inplace("mo_normalize_by_template", tensorV, param_id)


def DenormalizeByTemplate(tensorV, param_id):
    _inplace("mo_denormalize_by_template", tensorV, param_id)


# ---------------------------------------------------------------------------
# Distance layer (distance_layer.cc:8-80)

def _distance(name, tensorV, param_id, width):
    _check(tensorV, torch.float32, "tensorV", 3)
    V = _to_dev(tensorV)
    n = _nrows(V)
    out = torch.empty((n,) if width == 1 else (n, 3), dtype=torch.float32, device=V.device)
    _lib.call(name, _ptr(V), n, int(param_id), _ptr(out), _stream())
    return out.to(tensorV.device)


def DistanceFieldLoss_forward(tensorV, param_id):
    return _distance("mo_distance_loss_forward", tensorV, param_id, 1)


def DistanceFieldLoss_backward(tensorV, param_id):
    return _distance("mo_distance_loss_backward", tensorV, param_id, 3)


# ---------------------------------------------------------------------------
# Rigid (faces) and graph (edges) layers (rigid_layer.cc, graph_layer.cc)

def _store(name, tensorV, tensorI, width, param_id):
    _check(tensorV, torch.float32, "tensorV", 3)
    _check(tensorI, torch.int32, "index tensor", width)
    V = _to_dev(tensorV)
    I = _index_to_dev(tensorI)
    _lib.call(name, _ptr(V), _nrows(V), _ptr(I), _nrows(I), int(param_id), _stream())


def _edge_fwd(name, tensorV, tensorI, width, param_id):
    _check(tensorV, torch.float32, "tensorV", 3)
    _check(tensorI, torch.int32, "index tensor", width)
    V = _to_dev(tensorV)
    I = _index_to_dev(tensorI)
    nh = _nrows(I) * (3 if width == 3 else 1)
    out = torch.empty((nh, 3), dtype=torch.float32, device=V.device)
    _lib.call(name, _ptr(V), _nrows(V), _ptr(I), _nrows(I), int(param_id), _ptr(out), _stream())
    return out.to(tensorV.device)


def _edge_bwd(name, tensorV, tensorI, width, param_id):
    _check(tensorV, torch.float32, "tensorV", 3)
    _check(tensorI, torch.int32, "index tensor", width)
    V = _to_dev(tensorV)
    I = _index_to_dev(tensorI)
    out = torch.empty((_nrows(V), 3), dtype=torch.float32, device=V.device)
    _lib.call(name, _ptr(V), _nrows(V), _ptr(I), _nrows(I), int(param_id), _ptr(out), _stream())
    return out.to(tensorV.device)


def StoreRigidityInformation(tensorV, tensorF, param_id):
    _store("mo_rigidity_store", tensorV, tensorF, 3, param_id)


def RigidEdgeLoss_forward(tensorV, tensorF, param_id):
    return _edge_fwd("mo_rigid_loss_forward", tensorV, tensorF, 3, param_id)


def RigidEdgeLoss_backward(tensorV, tensorF, param_id):
    return _edge_bwd("mo_rigid_loss_backward", tensorV, tensorF, 3, param_id)


def StoreGraphInformation(tensorV, tensorE, param_id):
    _store("mo_graph_store", tensorV, tensorE, 2, param_id)


def GraphEdgeLoss_forward(tensorV, tensorE, param_id):
    return _edge_fwd("mo_graph_loss_forward", tensorV, tensorE, 2, param_id)


def GraphEdgeLoss_backward(tensorV, tensorE, param_id):
    return _edge_bwd("mo_graph_loss_backward", tensorV, tensorE, 2, param_id)


# ---------------------------------------------------------------------------
# Extras of this build (no pybind counterpart): template grid access.

def TemplateInfo(param_id):
    """(grid_resolution, scale, trans[3]) of a template."""
    import numpy as np
    N = C.c_int32()
    sc = C.c_double()
    tr = np.zeros(3)
    _lib.call("mo_template_info", int(param_id), C.byref(N), C.byref(sc), C.c_void_p(tr.ctypes.data))
    return int(N.value), float(sc.value), tr


def TemplateGrid(param_id):
    """(dist64, dist32, nearest) device tensors of shape [N,N,N] ([z][y][x])."""
    N, _, _ = TemplateInfo(param_id)
    dev = _device()
    d64 = torch.empty((N, N, N), dtype=torch.float64, device=dev)
    d32 = torch.empty((N, N, N), dtype=torch.float32, device=dev)
    nn = torch.empty((N, N, N), dtype=torch.int32, device=dev)
    _lib.call("mo_template_grid", int(param_id), _ptr(d64), _ptr(d32), _ptr(nn), _stream())
    return d64, d32, nn
