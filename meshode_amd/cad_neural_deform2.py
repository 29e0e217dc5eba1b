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
libmeshode; the flow's MLP runs in torch.  The reference evaluates the final
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


def deform(mesh1, mesh2, rigidity=0.1, device=torch.device("cuda"), niter=1000, log=print):
    """mesh1 / mesh2: LoadCadMesh bundles [V, F, E, V2G, GV, GE] (CPU
    tensors; GV is normalized in place, as the reference's layer does).
    Returns dict(V, F, losses [niter, 4], func, optimizer, param_id1,
    param_id2, GV1_deformed)."""
    V1, F1, E1, V2G1, GV1, GE1 = mesh1
    V2, F2, E2, V2G2, GV2, GE2 = mesh2
    graph_loss = GraphLoss2Layer(V1, F1, GV1, GE1, V2, F2, GV2, GE2, rigidity, device)
    param_id1 = graph_loss.param_id1
    param_id2 = graph_loss.param_id2
    reverse_loss = ReverseLossLayer()
    func = NeuralODE(device)
    optimizer = optim.Adam(func.parameters(), lr=1e-3)
    GV1_origin = GV1.clone().to(device)
    GV2_origin = GV2.clone().to(device)
    GV1_device = GV1.to(device)
    GV2_device = GV2.to(device)
    GE1_device = GE1.to(device)
    GE2_device = GE2.to(device)
    n1, n2 = GV1.shape[0], GV2.shape[0]
    losses = []
    for it in range(niter):
        optimizer.zero_grad()
        GV1_deformed = func.forward(GV1_device)
        GV2_deformed = func.inverse(GV2_device)
        loss1_forward = graph_loss(GV1_deformed, GE1_device, GV2_device, GE2_device, 0)
        loss1_backward = reverse_loss(GV1_deformed, GV2_origin, device)
        loss2_forward = graph_loss(GV1_device, GE1_device, GV2_deformed, GE2_device, 1)
        loss2_backward = reverse_loss(GV2_deformed, GV1_origin, device)
        loss = loss1_forward + loss1_backward + loss2_forward + loss2_backward
        loss.backward()
        optimizer.step()
        row = (loss1_forward.item(), loss1_backward.item(), loss2_forward.item(), loss2_backward.item())
        losses.append(row)
        if log is not None:
            log("iter=%d, loss1_forward=%.6f loss1_backward=%.6f loss2_forward=%.6f loss2_backward=%.6f"
                % (it, math.sqrt(row[0] / n1), math.sqrt(row[1] / n2), math.sqrt(row[2] / n2),
                   math.sqrt(row[3] / n1)))
    with torch.no_grad():
        GV1_deformed = func.forward(GV1_device)
        # the source mesh through the flow, in the source template's frame
        V1_copy = V1.clone()
        pyDeform.NormalizeByTemplate(V1_copy, param_id1.tolist())
        V1_origin = V1_copy.clone()
        V1_pushed = func.forward(V1_copy.to(device))
    src_to_src = torch.from_numpy(np.arange(V1_origin.shape[0]).astype("int32"))
    pyDeform.SolveLinear(V1_origin, F1, E1, src_to_src, V1_pushed, 1, 1)
    pyDeform.DenormalizeByTemplate(V1_origin, param_id2.tolist())
    return dict(V=V1_origin, F=F1, losses=np.array(losses, np.float64).reshape(-1, 4), func=func,
                optimizer=optimizer, param_id1=param_id1, param_id2=param_id2, GV1_deformed=GV1_deformed,
                V1_pushed=V1_pushed)


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
