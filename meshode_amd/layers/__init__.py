"""torch.autograd layers over the pyDeform drop-in, one module per
reference module (src/python/layers/*.py):

* rigid_loss_layer: RigidLossFunction, RigidLossLayer, Finalize
* graph_loss_layer: GraphLossFunction, GraphLossLayer
* graph_loss2_layer: GraphLoss2Function, GraphLoss2Layer, Finalize
* reverse_loss_layer: ReverseLossLayer
* neuralode_fast: ODEFunc, NeuralODE

The rigid and graph layers are re-exported here.
"""
from .graph_loss_layer import GraphLossFunction, GraphLossLayer  # noqa: F401
from .rigid_loss_layer import Finalize, RigidLossFunction, RigidLossLayer  # noqa: F401
