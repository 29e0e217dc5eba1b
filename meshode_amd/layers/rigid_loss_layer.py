"""RigidLossFunction / RigidLossLayer / Finalize: a restatement of
src/python/layers/rigid_loss_layer.py:7-44 over the pyDeform drop-in.  The
forward and backward passes call the hand-written HIP kernels through
pyDeform; tensors stay on the device they arrive on.
"""
from __future__ import annotations

import torch
from torch import nn
from torch.autograd import Function

from .. import pyDeform


class RigidLossFunction(Function):
    """loss = 0.5 sum D + 0.5 sum R (rigid_loss_layer.py:9-17); the backward
    adds the two hand-written gradients (rigid_loss_layer.py:19-27)."""

    @staticmethod
    def forward(ctx, src_V, src_F, param_id):
        pid = int(param_id.tolist())
        lossD = pyDeform.DistanceFieldLoss_forward(src_V, pid) * 0.5
        lossR = pyDeform.RigidEdgeLoss_forward(src_V, src_F, pid) * 0.5
        ctx.save_for_backward(src_V, src_F, param_id)
        return lossD.sum() + lossR.sum()

    @staticmethod
    def backward(ctx, grad_h):
        src_V, src_F, param_id = ctx.saved_tensors
        pid = int(param_id.tolist())
        gD = pyDeform.DistanceFieldLoss_backward(src_V, pid)
        gR = pyDeform.RigidEdgeLoss_backward(src_V, src_F, pid)
        return grad_h * (gD + gR), None, None


class RigidLossLayer(nn.Module):
    """rigid_loss_layer.py:30-41: template on the target (grid 64 by default),
    source normalized in place, rest offsets stored."""

    def __init__(self, src_V, src_F, tar_V, tar_F, grid_resolution=64):
        super().__init__()
        self.param_id = torch.tensor(pyDeform.InitializeDeformTemplate(tar_V, tar_F, 0, grid_resolution))
        pyDeform.NormalizeByTemplate(src_V, self.param_id.tolist())
        pyDeform.StoreRigidityInformation(src_V, src_F, self.param_id.tolist())

    def forward(self, src_V, src_F):
        return RigidLossFunction.apply(src_V, src_F, self.param_id)


def Finalize(src_V, param_id):
    """rigid_loss_layer.py:43-44."""
    pyDeform.DenormalizeByTemplate(src_V, int(param_id.tolist()))
