"""Levenberg-Marquardt deformation on MI355X (Deformer, src/lib/deformer.h:11-30).

``deform_batch`` runs Deformer::Deform (faces) / Deformer::DeformGraph
(edges) for many pairs in one persistent launch; ``deform_rot`` runs
Deformer::DeformWithRot (positions and per-vertex rotations); ``rigid_deform_batch`` runs
the whole rigid_deform pipeline (rigid_deform.cc:49-61: normalize target,
build its distance grid, move the source into the target frame, solve,
return the result in the target's original frame) device-resident.
"""
from __future__ import annotations

import ctypes as C

import numpy as np
import torch

from . import _lib


def default_options(**overrides) -> _lib.LMOptions:
    o = _lib.default_lm_options()
    for k, v in overrides.items():
        if not hasattr(o, k):
            raise AttributeError(f"unknown LM option {k!r}")
        setattr(o, k, v)
    return o


def _dev():
    if not torch.cuda.is_available():
        raise RuntimeError("deformation needs a ROCm GPU: torch.cuda.is_available() is False")
    return torch.device("cuda", torch.cuda.current_device())


def _as_dev(a, dtype):
    if isinstance(a, np.ndarray):
        a = torch.from_numpy(np.ascontiguousarray(a))
    return a.to(device=_dev(), dtype=dtype).contiguous()


def _arr(ctype, vals):
    return (ctype * len(vals))(*vals)


def _stream():
    return C.c_void_p(torch.cuda.current_stream().cuda_stream)


def deform_batch(param_ids, Vs, idxs, kinds, lambdas, options=None, no_alloc=False):
    """Deformer::Deform / DeformGraph for each pair.

    Vs: FP64 device tensors [nv,3] in the template's normalized frame,
    updated in place.  idxs: int32 faces [n,3] (kind 0) or edges [n,2]
    (kind 1).  no_alloc: solve through mo_deform_batch_noalloc (no device
    allocation; the batch's scratch must fit the library's fixed array).
    Returns a list of stats dicts."""
    n = len(param_ids)
    Vs = list(Vs)
    for V in Vs:
        if V.dtype != torch.float64 or not V.is_cuda or not V.is_contiguous():
            raise TypeError("deform_batch expects contiguous float64 device tensors")
    I = [_as_dev(i, torch.int32) for i in idxs]
    stats = (_lib.LMStats * n)()
    opt = options or default_options()
    _lib.call("mo_deform_batch_noalloc" if no_alloc else "mo_deform_batch", n,
              _arr(C.c_int32, [int(p) for p in param_ids]),
              _arr(C.c_void_p, [V.data_ptr() for V in Vs]),
              _arr(C.c_int64, [V.shape[0] for V in Vs]),
              _arr(C.c_void_p, [i.data_ptr() for i in I]),
              _arr(C.c_int64, [i.shape[0] for i in I]),
              _arr(C.c_int32, [int(k) for k in kinds]),
              _arr(C.c_double, [float(l) for l in lambdas]),
              C.byref(opt), stats, _stream())
    return [s.as_dict() for s in stats]


def deform_rot(param_id, V, F, lam=1.0, W=None, options=None):
    """Deformer::DeformWithRot (deformer.cc:94-171) on the GPU.

    V: FP64 device tensor [nv,3] in the template's normalized frame (rest
    shape in, solution out).  F: int32 faces.  W: optional FP64 device
    tensor [nv,3] of start rotations (zeros otherwise), updated in place.
    Returns (stats dict, rotations [nv,3] FP64 device tensor)."""
    if V.dtype != torch.float64 or not V.is_cuda or not V.is_contiguous() or V.dim() != 2 or V.shape[1] != 3:
        raise TypeError("deform_rot expects a contiguous [n,3] float64 device tensor")
    Fd = _as_dev(F, torch.int32).reshape(-1, 3)
    if W is None:
        W = torch.zeros_like(V)
    elif W.dtype != torch.float64 or not W.is_cuda or not W.is_contiguous() or W.shape != V.shape:
        raise TypeError("deform_rot: W must be a contiguous float64 device tensor shaped like V")
    st = _lib.LMStats()
    # options None: the library's DeformWithRot defaults (inner PCG tolerance 1e-6)
    _lib.call("mo_deform_rot", int(param_id), C.c_void_p(V.data_ptr()), V.shape[0], C.c_void_p(Fd.data_ptr()),
              Fd.shape[0], float(lam), C.byref(options) if options is not None else None, C.c_void_p(W.data_ptr()),
              C.byref(st), _stream())
    return st.as_dict(), W


def deform_rot_batch(param_ids, Vs, Fs, lambdas, options=None):
    """deform_rot for many independent problems in one launch.  Vs: FP64
    device tensors [nv,3] (normalized frames, updated in place); Fs: int32
    faces.  Returns (list of stats dicts, list of rotation tensors)."""
    n = len(param_ids)
    Vs = list(Vs)
    for V in Vs:
        if V.dtype != torch.float64 or not V.is_cuda or not V.is_contiguous() or V.dim() != 2 or V.shape[1] != 3:
            raise TypeError("deform_rot_batch expects contiguous [n,3] float64 device tensors")
    Fd = [_as_dev(f, torch.int32).reshape(-1, 3) for f in Fs]
    Ws = [torch.zeros_like(V) for V in Vs]
    stats = (_lib.LMStats * n)()
    _lib.call("mo_deform_rot_batch", n,
              _arr(C.c_int32, [int(p) for p in param_ids]),
              _arr(C.c_void_p, [V.data_ptr() for V in Vs]),
              _arr(C.c_int64, [V.shape[0] for V in Vs]),
              _arr(C.c_void_p, [f.data_ptr() for f in Fd]),
              _arr(C.c_int64, [f.shape[0] for f in Fd]),
              _arr(C.c_double, [float(l) for l in lambdas]),
              C.byref(options) if options is not None else None,
              _arr(C.c_void_p, [W.data_ptr() for W in Ws]), stats, _stream())
    return [s.as_dict() for s in stats], Ws


class PairBatch:
    """Device-resident inputs of a rigid_deform batch (FP64 vertices, int32 faces)."""

    def __init__(self, pairs, grid_resolution=64, lambdas=1.0):
        self.src = [_as_dev(p[0], torch.float64) for p in pairs]
        self.src_f = [_as_dev(p[1], torch.int32) for p in pairs]
        self.tgt = [_as_dev(p[2], torch.float64) for p in pairs]
        self.tgt_f = [_as_dev(p[3], torch.int32) for p in pairs]
        n = len(pairs)
        gr = grid_resolution if isinstance(grid_resolution, (list, tuple)) else [grid_resolution] * n
        lam = lambdas if isinstance(lambdas, (list, tuple)) else [lambdas] * n
        self.out = [torch.empty_like(s) for s in self.src]
        self._args = (
            n,
            _arr(C.c_void_p, [t.data_ptr() for t in self.src]),
            _arr(C.c_int64, [t.shape[0] for t in self.src]),
            _arr(C.c_void_p, [t.data_ptr() for t in self.src_f]),
            _arr(C.c_int64, [t.shape[0] for t in self.src_f]),
            _arr(C.c_void_p, [t.data_ptr() for t in self.tgt]),
            _arr(C.c_int64, [t.shape[0] for t in self.tgt]),
            _arr(C.c_void_p, [t.data_ptr() for t in self.tgt_f]),
            _arr(C.c_int64, [t.shape[0] for t in self.tgt_f]),
            _arr(C.c_int32, [int(g) for g in gr]),
            _arr(C.c_double, [float(l) for l in lam]),
        )
        self._out_ptrs = _arr(C.c_void_p, [t.data_ptr() for t in self.out])
        self.n = n

    def run(self, options=None):
        opt = options or default_options()
        stats = (_lib.LMStats * self.n)()
        _lib.call("mo_rigid_deform_batch", *self._args, C.byref(opt), self._out_ptrs, stats, _stream())
        return [s.as_dict() for s in stats]

    def run_async(self, options=None, after=None):
        """mo_rigid_deform_batch_async: enqueue the batch on the current
        stream and return once its solver launches are queued; wait()
        returns its stats.  after: another PairBatch with a batch in flight
        (on another stream) whose solver launches this batch's grid builds
        follow.  One batch of this object in flight at a time (its outputs
        are self.out)."""
        if getattr(self, "_pending", None) is not None:
            raise RuntimeError("a batch of this PairBatch is still in flight: call wait() first")
        prev = None
        if after is not None:
            prev = getattr(after, "_pending", None)
            if prev is None:
                raise RuntimeError("`after` has no batch in flight")
        opt = options or default_options()
        h = C.c_void_p()
        _lib.call("mo_rigid_deform_batch_async", *self._args, C.byref(opt), self._out_ptrs, _stream(), prev,
                  C.byref(h))
        self._pending = h

    def wait(self):
        """mo_batch_wait on the batch in flight: waits for it and returns its
        stats (as run does)."""
        h = getattr(self, "_pending", None)
        if h is None:
            raise RuntimeError("no batch in flight: call run_async() first")
        self._pending = None
        stats = (_lib.LMStats * self.n)()
        _lib.call("mo_batch_wait", h, stats)
        return [s.as_dict() for s in stats]

    def capture(self, options=None):
        """mo_rigid_deform_batch_graph: solve the batch once (returns its
        stats, like run) and capture the whole launch chain into a HIP graph
        that replay() re-runs on the current contents of the same inputs."""
        self.release()
        opt = options or default_options()
        stats = (_lib.LMStats * self.n)()
        g = C.c_void_p()
        _lib.call("mo_rigid_deform_batch_graph", *self._args, C.byref(opt), self._out_ptrs, stats, _stream(),
                  C.byref(g))
        self._graph = g
        return [s.as_dict() for s in stats]

    def replay(self):
        """mo_batch_graph_launch: one replay of the captured chain (waits for
        it); returns the replay's stats."""
        if getattr(self, "_graph", None) is None:
            raise RuntimeError("no captured graph: call capture() first")
        stats = (_lib.LMStats * self.n)()
        _lib.call("mo_batch_graph_launch", self._graph, stats, _stream())
        return [s.as_dict() for s in stats]

    def release(self):
        """mo_batch_graph_destroy of the captured graph, if any."""
        g = getattr(self, "_graph", None)
        if g is not None:
            self._graph = None
            _lib.call("mo_batch_graph_destroy", g)


def rigid_deform_batch(pairs, grid_resolution=64, lambdas=1.0, options=None):
    """Run the rigid_deform pipeline on [(src_V, src_F, tgt_V, tgt_F), ...].
    Returns (list of deformed source vertex tensors in the target frame, stats)."""
    b = PairBatch(pairs, grid_resolution, lambdas)
    st = b.run(options)
    return b.out, st


def template_f64(V, F, grid_resolution):
    """mo_template_create_f64: template from FP64 vertices (the CLI path,
    Mesh::ReadOBJ keeps doubles).  Returns the param_id."""
    Vd = _as_dev(V, torch.float64)
    Fd = _as_dev(F, torch.int32)
    pid = C.c_int32(-1)
    _lib.call("mo_template_create_f64", C.c_void_p(Vd.data_ptr()), Vd.shape[0], C.c_void_p(Fd.data_ptr()),
              Fd.shape[0], 0, int(grid_resolution), _stream(), C.byref(pid))
    return int(pid.value)


def template_rebuild_f64(param_id, V, F):
    """mo_template_rebuild_f64: rebuild an existing template in place for a
    new FP64 target (same grid resolution; no grid allocation or free)."""
    Vd = _as_dev(V, torch.float64)
    Fd = _as_dev(F, torch.int32)
    _lib.call("mo_template_rebuild_f64", int(param_id), C.c_void_p(Vd.data_ptr()), Vd.shape[0],
              C.c_void_p(Fd.data_ptr()), Fd.shape[0], _stream())
