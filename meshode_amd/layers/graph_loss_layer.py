"""GraphLossFunction / GraphLossLayer: a restatement of
src/python/layers/graph_loss_layer.py:9-61 over the pyDeform drop-in (HIP
kernels; tensors stay on the device they arrive on).
"""
from __future__ import annotations

import torch
from torch import nn
from torch.autograd import Function

from .. import pyDeform


class GraphLossFunction(Function):
    """graph_loss_layer.py:9-43: loss = 0.5 sum D + 0.5 rigidity^2 sum R; the
    distance gradient is masked where 0.5 d^2 >= 0.5 * 0.03^2 (the forward is
    not masked, as in the reference)."""

    @staticmethod
    def forward(ctx, src_V, src_E, rigidity2, param_id):
        pid = int(param_id.tolist())
        lossD = pyDeform.DistanceFieldLoss_forward(src_V, pid) * 0.5
        lossR = pyDeform.GraphEdgeLoss_forward(src_V, src_E, pid) * 0.5
        mask_D = lossD < (0.5 * 0.03 * 0.03)
        ctx.save_for_backward(src_V, src_E, rigidity2, param_id, mask_D)
        return lossD.sum() + lossR.sum() * rigidity2.tolist()

    @staticmethod
    def backward(ctx, grad_h):
        src_V, src_E, rigidity2, param_id, mask_D = ctx.saved_tensors
        pid = int(param_id.tolist())
        mask = mask_D.view(-1, 1).expand(-1, 3)
        gD = pyDeform.DistanceFieldLoss_backward(src_V, pid) * mask
        gR = pyDeform.GraphEdgeLoss_backward(src_V, src_E, pid)
        return grad_h * (gD + gR * rigidity2.tolist()), None, None, None


class GraphLossLayer(nn.Module):
    """graph_loss_layer.py:46-57."""

    def __init__(self, src_V, src_E, tar_V, tar_F, rigidity, grid_resolution=64):
        super().__init__()
        self.param_id = torch.tensor(pyDeform.InitializeDeformTemplate(tar_V, tar_F, 0, grid_resolution))
        pyDeform.NormalizeByTemplate(src_V, self.param_id.tolist())
        pyDeform.StoreGraphInformation(src_V, src_E, self.param_id.tolist())
        self.rigidity2 = torch.tensor(rigidity * rigidity)

    def forward(self, src_V, src_E):
        return GraphLossFunction.apply(src_V, src_E, self.rigidity2, self.param_id)
