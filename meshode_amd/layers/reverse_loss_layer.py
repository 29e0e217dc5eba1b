"""ReverseLossLayer: a restatement of src/python/layers/reverse_loss_layer.py:10-26.

For every target point the nearest source point (the reference builds a
scipy cKDTree on the source and queries k = 1) is found by the
``mo_nearest_f32`` kernel; loss = 1/2 sum |src[nn] - tar|^2 in FP32 torch
arithmetic, so autograd differentiates it with respect to the source
through the gather, as in the reference.  Squared distances are compared
in FP64 from the FP32 coordinates; among equal distances the lowest source
index wins (cKDTree leaves ties unspecified).
"""
from __future__ import annotations

import ctypes as C

import torch
from torch import nn

from .. import _lib, pyDeform


def nearest(src_V, tar_V, return_dist2=False):
    """Index (int64, on tar_V's device) of the nearest src_V point for each
    tar_V point; optionally the FP64 squared distances."""
    src = pyDeform._to_dev(src_V.detach().reshape(-1, 3).to(torch.float32))
    tar = pyDeform._to_dev(tar_V.detach().reshape(-1, 3).to(torch.float32))
    if src.shape[0] == 0:
        raise ValueError("nearest: empty source point set")
    idx = torch.empty(tar.shape[0], dtype=torch.int32, device=src.device)
    d2 = torch.empty(tar.shape[0], dtype=torch.float64, device=src.device) if return_dist2 else None
    _lib.call("mo_nearest_f32", C.c_void_p(src.data_ptr()), src.shape[0], C.c_void_p(tar.data_ptr()), tar.shape[0],
              C.c_void_p(idx.data_ptr()), C.c_void_p(d2.data_ptr() if d2 is not None else 0), pyDeform._stream())
    idx = idx.to(tar_V.device, torch.int64)
    if return_dist2:
        return idx, d2.to(tar_V.device)
    return idx


class ReverseLossLayer(nn.Module):
    def __init__(self):
        super().__init__()

    def forward(self, src_V, tar_V, device=torch.device("cpu")):
        ii = nearest(src_V, tar_V).to(src_V.device)
        loss = src_V[ii] - tar_V.to(device)
        return 0.5 * (loss * loss).sum()
