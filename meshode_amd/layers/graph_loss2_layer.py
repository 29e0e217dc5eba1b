"""GraphLoss2Function / GraphLoss2Layer / Finalize: a restatement of
src/python/layers/graph_loss2_layer.py:9-77 over the pyDeform drop-in.

Two templates, one per mesh; each graph is normalized by and stores its
rest offsets in its OWN template (graph_loss2_layer.py:53-62), while its
distance term is looked up in the OTHER mesh's grid (:18-19, :37-38), as
the reference does.  Unlike GraphLossFunction, the distance gradient is not
masked.  The kernels run on the device the tensors are on; the reference's
CPU copies and ``.to(device)`` moves are not needed.
"""
from __future__ import annotations

import torch
from torch import nn
from torch.autograd import Function

from .. import pyDeform


class GraphLoss2Function(Function):
    """loss = 0.5 sum D(V1; grid2) + 0.5 rigidity^2 sum R(V1, E1; offsets1)
    (graph_loss2_layer.py:11-26); backward grad_h (dD + dR rigidity^2)
    (:28-44)."""

    @staticmethod
    def forward(ctx, V1, E1, rigidity2, param_id1, param_id2):
        pid1 = int(param_id1.tolist())
        pid2 = int(param_id2.tolist())
        lossD1 = pyDeform.DistanceFieldLoss_forward(V1, pid2) * 0.5
        lossR1 = pyDeform.GraphEdgeLoss_forward(V1, E1, pid1) * 0.5
        ctx.save_for_backward(V1, E1, rigidity2, param_id1, param_id2)
        return lossD1.sum() + lossR1.sum() * rigidity2.tolist()

    @staticmethod
    def backward(ctx, grad_h):
        V1, E1, rigidity2, param_id1, param_id2 = ctx.saved_tensors
        pid1 = int(param_id1.tolist())
        pid2 = int(param_id2.tolist())
        gD = pyDeform.DistanceFieldLoss_backward(V1, pid2)
        gR = pyDeform.GraphEdgeLoss_backward(V1, E1, pid1)
        return grad_h * (gD + gR * rigidity2.tolist()), None, None, None, None


class GraphLoss2Layer(nn.Module):
    """graph_loss2_layer.py:47-71: templates on (V1, F1) and (V2, F2) (grid
    64), graph_V1 / graph_V2 normalized in place by their own template and
    their edge offsets stored there."""

    def __init__(self, V1, F1, graph_V1, graph_E1, V2, F2, graph_V2, graph_E2, rigidity,
                 d=torch.device("cpu"), grid_resolution=64):
        super().__init__()
        self.device = d
        self.param_id1 = torch.tensor(pyDeform.InitializeDeformTemplate(V1, F1, 0, grid_resolution))
        self.param_id2 = torch.tensor(pyDeform.InitializeDeformTemplate(V2, F2, 0, grid_resolution))
        pyDeform.NormalizeByTemplate(graph_V1, self.param_id1.tolist())
        pyDeform.NormalizeByTemplate(graph_V2, self.param_id2.tolist())
        pyDeform.StoreGraphInformation(graph_V1, graph_E1, self.param_id1.tolist())
        pyDeform.StoreGraphInformation(graph_V2, graph_E2, self.param_id2.tolist())
        self.rigidity2 = torch.tensor(rigidity * rigidity)

    def forward(self, V1, E1, V2, E2, direction):
        if direction == 0:
            return GraphLoss2Function.apply(V1, E1, self.rigidity2, self.param_id1, self.param_id2)
        return GraphLoss2Function.apply(V2, E2, self.rigidity2, self.param_id2, self.param_id1)


def Finalize(src_V, src_F, src_E, src_to_graph, graph_V, rigidity, param_id):
    """graph_loss2_layer.py:73-77: normalize, LinearEstimation
    (SolveLinear with_rot = 0), denormalize, in place."""
    pyDeform.NormalizeByTemplate(src_V, param_id.tolist())
    pyDeform.SolveLinear(src_V, src_F, src_E, src_to_graph, graph_V, rigidity)
    pyDeform.DenormalizeByTemplate(src_V, param_id.tolist())
