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


def _alone(r):
    out, st = deform.rigid_deform_batch([(r.src_V, r.src_F, r.tgt_V, r.tgt_F)], grid_resolution=r.grid, lambdas=r.lam)
    return out[0].cpu().numpy(), st[0]


def test_served_100k_request_at_256_equals_alone(gpu):
    """Config 5's largest request: a 100,000-vertex source against its
    target at 256^3 (a 64 MB FP32 grid), served together with small
    requests on other targets: its result, and theirs, equal each pair
    solved alone."""
    print("[serve] case=100k source, 256^3 grid, in a mix", flush=True)
    big = synth.box_union_pair(81, n_verts=100000)
    assert len(big[0]) == 100000
    small = [synth.box_union_pair(82 + i, n_boxes=5, n_verts=3000) for i in range(3)]
    reqs = [serve.DeformRequest(big[0], big[1], "big", big[2], big[3], grid=256, lam=1.0)]
    reqs += [serve.DeformRequest(p[0], p[1], ("s", i), p[2], p[3], grid=64, lam=0.8) for i, p in enumerate(small)]
    srv = serve.DeformServer(mem_budget_bytes=8e9)
    try:
        futs = [srv.submit(r) for r in reqs]
        res = [f.result(timeout=180) for f in futs]
    finally:
        srv.close()
    assert res[0].stats["solver_path_name"] == "lds_team"
    for r, out in zip(reqs, res):
        x, st = _alone(r)
        assert np.array_equal(out.V, x)
        assert out.stats["final_cost"] == st["final_cost"]
    print(f"[serve] 100k at 256^3: latency {res[0].latency_s * 1e3:.0f} ms, {res[0].stats['iterations']} LM "
          f"iterations", flush=True)


def test_cache_eviction_under_small_budget(gpu):
    """Distinct targets through a budget that holds two 64^3 templates: the
    least recently used idle template is recycled in place for each new
    target (mo_template_rebuild_f64), a cached target is reused, the
    resident bytes never exceed the budget, and every result equals the
    pair solved alone."""
    print("[serve] case=LRU eviction with distinct targets", flush=True)
    pairs = [synth.box_union_pair(91 + i, n_boxes=4, n_verts=1000) for i in range(4)]
    one = serve.template_bytes(64)
    solver = max(serve.solver_bytes(len(p[0]), len(p[1])) for p in pairs)
    srv = serve.DeformServer(mem_budget_bytes=2 * one + solver + 1, small_workers=1)
    order = [0, 1, 0, 2, 0, 3, 1]
    hits = {2, 4}
    try:
        for step, k in enumerate(order):
            p = pairs[k]
            r = serve.DeformRequest(p[0], p[1], ("tgt", k), p[2], p[3], grid=64)
            out = srv.submit(r).result(timeout=120)
            keys, used, budget = srv.cache_state()
            assert used <= budget and len(keys) <= 2, (step, keys)
            assert keys[-1] == (("tgt", k), 64)
            x, st = _alone(r)
            assert np.array_equal(out.V, x), step
            assert out.template_built == (step not in hits), step
            print(f"[serve] step {step}: target {k}, resident {[kk[0][1] for kk in keys]}, built "
                  f"{out.template_built}", flush=True)
        # 0 and 1 built new; 0 hit; 2 recycles 1's slot; 0 hit; 3 recycles 2's; 1 recycles 0's
        assert srv.builds == 5 and srv.rebuilds == 3
    finally:
        srv.close()
