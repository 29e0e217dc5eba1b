"""NeuralODE bidirectional CAD deformation: a restatement of
src/python/cad_neural_deform2.py:1-108 on the MI355X drop-in.

    python -m meshode_amd.cad_neural_deform2 --source cad-source.obj \\
        --target cad-target.obj --output cad-output.obj [--rigidity 0.1] \\
        [--device cuda] [--save_path cad-output.ckpt] [--niter 1000]

Kept from the reference: both meshes through LoadCadMesh (:39-40); one
GraphLoss2Layer over the two graphs (:42); the NeuralODE flow (ODEFunc,
one RK4 step) trained with Adam (lr 1e-3) on
    graph_loss(forward(GV1), dir 0) + reverse(forward(GV1) -> GV2)
  + graph_loss(inverse(GV2), dir 1) + reverse(inverse(GV2) -> GV1)
(:55-80), printing the four root-mean losses every iteration (:82-89); the
checkpoint {'func', 'optim'} (:93-94); the source mesh pushed through the
flow in the source template's frame, then SolveLinear(with_rot = 1,
rigidity 1) towards the pushed positions, denormalized by the TARGET's
template and saved (:96-108).  ``--niter`` (default 1000, the reference's
constant) is added.

Everything runs on the device: the distance and graph-edge kernels, the
nearest-neighbour queries and LinearEstimationWithRot are HIP kernels in
libmeshode, and so is the flow: one fused RK4 step (the four stages of the
MLP, and its backward) per direction.  The reference evaluates the final
flow on the CPU; here it stays on the device.
"""
from __future__ import annotations

import argparse
import math

import numpy as np
import torch
import torch.optim as optim

from . import pyDeform
from .layers.graph_loss2_layer import GraphLoss2Layer
from .layers.neuralode_fast import NeuralODE
from .layers.reverse_loss_layer import ReverseLossLayer


class NeuralDeform2:
    """The driver's state: layers, flow, optimizer and device copies of the
    graphs (cad_neural_deform2.py:39-63).  ``step`` runs one iteration of the
    bidirectional loop (:68-80) and returns its four losses as a device
    tensor; ``finalize`` runs :96-108 and returns the deformed source
    vertices (float32, CPU) in the target's original frame."""

    def __init__(self, mesh1, mesh2, rigidity=0.1, device=torch.device("cuda")):
        self.mesh1, self.mesh2 = mesh1, mesh2
        V1, F1, E1, V2G1, GV1, GE1 = mesh1
        V2, F2, E2, V2G2, GV2, GE2 = mesh2
        self.device = device
        self.graph_loss = GraphLoss2Layer(V1, F1, GV1, GE1, V2, F2, GV2, GE2, rigidity, device)
        self.param_id1 = self.graph_loss.param_id1
        self.param_id2 = self.graph_loss.param_id2
        self.reverse_loss = ReverseLossLayer()
        self.func = NeuralODE(device)
        # Adam(lr = 1e-3) as the reference; on the GPU torch's fused form (one
        # kernel for all parameters, the same update formula)
        self.optimizer = optim.Adam(self.func.parameters(), lr=1e-3, fused=device.type == "cuda")
        self.GV1_origin = GV1.clone().to(device)
        self.GV2_origin = GV2.clone().to(device)
        self.GV1_device = GV1.to(device)
        self.GV2_device = GV2.to(device)
        self.GE1_device = GE1.to(device)
        self.GE2_device = GE2.to(device)

    def step(self):
        self.optimizer.zero_grad()
        GV1_deformed = self.func.forward(self.GV1_device)
        GV2_deformed = self.func.inverse(self.GV2_device)
        loss1_forward = self.graph_loss(GV1_deformed, self.GE1_device, self.GV2_device, self.GE2_device, 0)
        loss1_backward = self.reverse_loss(GV1_deformed, self.GV2_origin, self.device)
        loss2_forward = self.graph_loss(self.GV1_device, self.GE1_device, GV2_deformed, self.GE2_device, 1)
        loss2_backward = self.reverse_loss(GV2_deformed, self.GV1_origin, self.device)
        loss = loss1_forward + loss1_backward + loss2_forward + loss2_backward
        loss.backward()
        self.optimizer.step()
        return torch.stack([loss1_forward, loss1_backward, loss2_forward, loss2_backward]).detach()

    def finalize(self):
        V1, F1, E1 = self.mesh1[0], self.mesh1[1], self.mesh1[2]
        with torch.no_grad():
            self.GV1_deformed = self.func.forward(self.GV1_device)
            # the source mesh through the flow, in the source template's frame
            V1_copy = V1.clone()
            pyDeform.NormalizeByTemplate(V1_copy, self.param_id1.tolist())
            V1_origin = V1_copy.clone()
            self.V1_pushed = self.func.forward(V1_copy.to(self.device))
        src_to_src = torch.from_numpy(np.arange(V1_origin.shape[0]).astype("int32"))
        pyDeform.SolveLinear(V1_origin, F1, E1, src_to_src, self.V1_pushed, 1, 1)
        pyDeform.DenormalizeByTemplate(V1_origin, self.param_id2.tolist())
        return V1_origin


def deform(mesh1, mesh2, rigidity=0.1, device=torch.device("cuda"), niter=1000, log=print):
    """mesh1 / mesh2: LoadCadMesh bundles [V, F, E, V2G, GV, GE] (CPU
    tensors; GV is normalized in place, as the reference's layer does).
    Returns dict(V, F, losses [niter, 4], func, optimizer, param_id1,
    param_id2, GV1_deformed, V1_pushed)."""
    d = NeuralDeform2(mesh1, mesh2, rigidity, device)
    n1, n2 = mesh1[4].shape[0], mesh2[4].shape[0]
    losses = []
    for it in range(niter):
        row = d.step().tolist()
        losses.append(row)
        if log is not None:
            log("iter=%d, loss1_forward=%.6f loss1_backward=%.6f loss2_forward=%.6f loss2_backward=%.6f"
                % (it, math.sqrt(row[0] / n1), math.sqrt(row[1] / n2), math.sqrt(row[2] / n2),
                   math.sqrt(row[3] / n1)))
    V = d.finalize()
    return dict(V=V, F=mesh1[1], losses=np.array(losses, np.float64).reshape(-1, 4), func=d.func,
                optimizer=d.optimizer, param_id1=d.param_id1, param_id2=d.param_id2,
                GV1_deformed=d.GV1_deformed, V1_pushed=d.V1_pushed)


def main(argv=None):
    parser = argparse.ArgumentParser(description="Rigid Deformation.")
    parser.add_argument("--source", default="../data/cad-source.obj")
    parser.add_argument("--target", default="../data/cad-target.obj")
    parser.add_argument("--output", default="./cad-output.obj")
    parser.add_argument("--rigidity", default="0.1")
    parser.add_argument("--device", default="cuda")
    parser.add_argument("--save_path", default="./cad-output.ckpt")
    parser.add_argument("--niter", type=int, default=1000)
    args = parser.parse_args(argv)
    device = torch.device(args.device)
    mesh1 = pyDeform.LoadCadMesh(args.source)
    mesh2 = pyDeform.LoadCadMesh(args.target)
    r = deform(mesh1, mesh2, float(args.rigidity), device, args.niter)
    if args.save_path != "":
        torch.save({"func": r["func"], "optim": r["optimizer"]}, args.save_path)
    pyDeform.SaveMesh(args.output, r["V"], r["F"])
    return 0


if __name__ == "__main__":
    raise SystemExit(main())
