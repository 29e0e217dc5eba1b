"""Continuous-batching server: every request's result is bit-identical to the
same pair solved alone, whatever else shares its batch; cached targets are
reused; admission respects the memory budget."""
import numpy as np
import pytest
import torch

from meshode_amd import deform, serve, synth

pytestmark = pytest.mark.gpu


def test_results_same_alone_and_in_the_mix(gpu):
    print("[serve] case=mixed_6_requests", flush=True)
    pairs = {
        "boxA": synth.box_union_pair(51, n_boxes=6, n_verts=2000),
        "sph": synth.displaced_sphere_pair(52, src_level=2, tgt_level=3),
        "boxB": synth.box_union_pair(53, n_boxes=4, n_verts=1000),
    }
    reqs = []
    for i, key in enumerate(["boxA", "sph", "boxB", "boxA", "sph", "boxA"]):
        sV, sF, tV, tF = pairs[key]
        src = sV * (1.0 + 0.01 * i)          # distinct sources, shared targets
        reqs.append(serve.DeformRequest(src, sF, key, tV, tF, grid=32, lam=1.0 + 0.25 * (i % 2)))
    srv = serve.DeformServer(mem_budget_bytes=8e9, max_batch=16, workers=2)
    try:
        futs = [srv.submit(r) for r in reqs]
        res = [f.result(timeout=120) for f in futs]
    finally:
        srv.close()
    assert srv.completed == len(reqs)
    for r, out in zip(reqs, res):
        alone, st = deform.rigid_deform_batch([(r.src_V, r.src_F, r.tgt_V, r.tgt_F)], grid_resolution=r.grid,
                                              lambdas=r.lam)
        assert np.array_equal(out.V, alone[0].cpu().numpy())
        assert out.stats["final_cost"] == st[0]["final_cost"]
        assert out.latency_s > 0


def test_team_pairs_same_alone_and_in_the_mix(gpu):
    """Requests above 8192 source vertices (multi-workgroup solves, one
    cooperative launch at a time across the two worker streams) mixed with
    small ones: each result equals the pair solved alone."""
    print("[serve] case=team_pairs", flush=True)
    pairs = {
        "big": synth.box_union_pair(71, n_boxes=8, n_verts=12000),
        "sph5": synth.displaced_sphere_pair(72, src_level=5, tgt_level=4),
        "small": synth.box_union_pair(73, n_boxes=4, n_verts=1000),
    }
    reqs = []
    for i, key in enumerate(["big", "small", "sph5", "big", "small", "sph5", "big"]):
        sV, sF, tV, tF = pairs[key]
        reqs.append(serve.DeformRequest(sV * (1.0 + 0.01 * i), sF, key, tV, tF, grid=32, lam=1.0))
    srv = serve.DeformServer(mem_budget_bytes=16e9, max_batch=4, workers=2)
    try:
        futs = [srv.submit(r) for r in reqs]
        res = [f.result(timeout=120) for f in futs]
    finally:
        srv.close()
    for r, out in zip(reqs, res):
        alone, st = deform.rigid_deform_batch([(r.src_V, r.src_F, r.tgt_V, r.tgt_F)], grid_resolution=r.grid,
                                              lambdas=r.lam)
        assert np.array_equal(out.V, alone[0].cpu().numpy())
        assert out.stats["final_cost"] == st[0]["final_cost"]


def test_admission_splits_batches(gpu):
    print("[serve] case=budget", flush=True)
    sV, sF, tV, tF = synth.box_union_pair(61, n_boxes=4, n_verts=1000)
    one = serve.solver_bytes(len(sV), len(sF)) + serve.template_bytes(32, len(tV), len(tF))
    srv = serve.DeformServer(mem_budget_bytes=1.5 * one, max_batch=64, workers=1)
    try:
        futs = [srv.submit(serve.DeformRequest(sV, sF, ("t", i), tV, tF, grid=32)) for i in range(4)]
        res = [f.result(timeout=120) for f in futs]
    finally:
        srv.close()
    assert all(r.batch_size == 1 for r in res)       # the budget admits one pair at a time
    assert all(np.array_equal(res[0].V, r.V) for r in res)
