"""Asynchronous rigid_deform batches (mo_rigid_deform_batch_async /
mo_batch_wait; the reference runs one ceres::Solve per process,
rigid_deform.cc:12-63).

Bar: an asynchronous batch runs the same kernels on the same inputs as
mo_rigid_deform_batch, so its outputs and stats equal a synchronous call's
bit for bit, also with two batches in flight on two streams (the
synchronous path's energy parity with the oracle is pinned in
test_gpu_lm.py and test_gpu_configs.py)."""
import pytest
import torch

from meshode_amd import _lib, deform, synth

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, flush=True)


def _pairs(base, n):
    return [synth.box_union_pair(base + k, n_verts=1200 + 500 * k) for k in range(n)]


def _same_stats(a, b):
    for x, y in zip(a, b):
        for k in ("iterations", "final_cost", "initial_cost", "pcg_iterations", "termination", "solver_path",
                  "workgroups"):
            assert x[k] == y[k], (k, x[k], y[k])


def test_async_equals_sync(gpu):
    _say("[async] case=6 pairs, grids 32/48, one batch in flight")
    pairs = _pairs(80, 6)
    grids = [32, 48] * 3
    opt = deform.default_options(max_num_iterations=20)
    ref = deform.PairBatch(pairs, grid_resolution=grids)
    st_ref = ref.run(opt)
    b = deform.PairBatch(pairs, grid_resolution=grids)
    b.run_async(opt)
    st = b.wait()
    for k in range(len(pairs)):
        assert torch.equal(b.out[k], ref.out[k]), k
    _same_stats(st, st_ref)
    with pytest.raises(RuntimeError, match="no batch in flight"):
        b.wait()


def test_two_batches_in_flight_on_two_streams(gpu):
    """Batches alternate between two streams with two in flight, as a
    pipelined caller runs them: every batch equals its synchronous run."""
    _say("[async] case=two streams, 4 batches of 5 pairs at 48^3")
    sets = [_pairs(90 + 10 * j, 5) for j in range(4)]
    opt = deform.default_options(max_num_iterations=20)
    refs = []
    for ps in sets:
        r = deform.PairBatch(ps, grid_resolution=48)
        refs.append((r, r.run(opt)))
    batches = [deform.PairBatch(ps, grid_resolution=48) for ps in sets]
    streams = [torch.cuda.Stream(), torch.cuda.Stream()]
    torch.cuda.synchronize()
    stats = [None] * len(batches)
    prev = None
    for j, b in enumerate(batches):
        with torch.cuda.stream(streams[j % 2]):
            # odd batches follow the previous batch's solver launches
            b.run_async(opt, after=batches[prev] if (prev is not None and j % 2) else None)
        if prev is not None:
            stats[prev] = batches[prev].wait()
        prev = j
    stats[prev] = batches[prev].wait()
    for j, (b, (r, st_r)) in enumerate(zip(batches, refs)):
        for k in range(len(sets[j])):
            assert torch.equal(b.out[k], r.out[k]), (j, k)
        _same_stats(stats[j], st_r)
        _say(f"[async] batch {j} equal")


def test_async_team_pair_completes_in_call(gpu):
    """A pair above one workgroup takes a cooperative team launch: the call
    completes the batch itself and wait returns the same stats."""
    _say("[async] case=12000-vertex pair (team) + small pair")
    pairs = [synth.box_union_pair(73, n_verts=12000), synth.box_union_pair(74, n_verts=900)]
    opt = deform.default_options(max_num_iterations=3)
    ref = deform.PairBatch(pairs, grid_resolution=32)
    st_ref = ref.run(opt)
    assert st_ref[0]["solver_path_name"] in ("lds_team", "team")
    b = deform.PairBatch(pairs, grid_resolution=32)
    b.run_async(opt)
    st = b.wait()
    for k in range(2):
        assert torch.equal(b.out[k], ref.out[k]), k
    _same_stats(st, st_ref)


def test_async_errors(gpu):
    """timing on is refused before anything runs; a target face index out of
    range is skipped by the grid build and reported by the wait, as the
    synchronous call reports it; a second batch on one PairBatch is refused
    while the first is in flight."""
    _say("[async] case=argument and face-index errors")
    pairs = [synth.box_union_pair(75, n_boxes=4, n_verts=800)]
    b = deform.PairBatch(pairs, grid_resolution=32)
    with pytest.raises(_lib.MoError, match="timing"):
        b.run_async(deform.default_options(timing=1))
    opt = deform.default_options(max_num_iterations=5)
    b.run_async(opt)
    with pytest.raises(RuntimeError, match="still in flight"):
        b.run_async(opt)
    c = deform.PairBatch(pairs, grid_resolution=32)
    d = deform.PairBatch(pairs, grid_resolution=32)
    with pytest.raises(RuntimeError, match="`after` has no batch in flight"):
        d.run_async(opt, after=c)
    b.wait()
    b.tgt_f[0][0, 0] = b.tgt[0].shape[0] + 5
    with pytest.raises(_lib.MoError, match="index out of range"):
        b.run(opt)
    b.run_async(opt)
    with pytest.raises(_lib.MoError, match="index out of range"):
        b.wait()
