"""HIP-graph replay of the rigid_deform batch pipeline
(mo_rigid_deform_batch_graph / mo_batch_graph_launch; the reference runs one
ceres::Solve per process, rigid_deform.cc:12-63).

Bar: the captured graph runs the same kernels on the same buffers as
mo_rigid_deform_batch, so every replay equals an eager call on the same
inputs bit for bit (the eager path's energy parity with the oracle is
pinned in test_gpu_lm.py and test_gpu_configs.py).  Each replay is checked,
with eager work and synchronizations between the replays."""
import math

import pytest
import torch

from meshode_amd import _lib, deform, synth

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, flush=True)


def _pairs():
    ps = [synth.box_union_pair(60 + k, n_verts=1500 + 700 * k) for k in range(6)]
    ps.append(synth.displaced_sphere_pair(5, src_level=3, tgt_level=4))
    return ps


GRIDS = [32, 48, 64, 32, 48, 64, 32]


def _move(t, k):
    """A rigid motion of the k-th source, in place."""
    a = 0.05 * (k + 1)
    R = torch.tensor([[math.cos(a), -math.sin(a), 0.0], [math.sin(a), math.cos(a), 0.0], [0.0, 0.0, 1.0]],
                     dtype=torch.float64, device=t.device)
    t.copy_(t @ R.T + torch.tensor([0.01 * k, -0.02, 0.015], dtype=torch.float64, device=t.device))


def test_graph_replays_equal_eager(gpu):
    _say("[graph] case=7 pairs, grids 32/48/64, 3 replays + moved sources")
    pairs = _pairs()
    opt = deform.default_options(max_num_iterations=20)
    eager = deform.PairBatch(pairs, grid_resolution=GRIDS)
    st_e = eager.run(opt)
    ref = [o.clone() for o in eager.out]
    g = deform.PairBatch(pairs, grid_resolution=GRIDS)
    st_c = g.capture(opt)
    assert all(s["solver_path_name"] == "lds" for s in st_c)
    for k in range(len(pairs)):
        assert torch.equal(g.out[k], ref[k]), k
    for rep in range(3):
        for o in g.out:
            o.zero_()
        torch.cuda.synchronize()
        st_r = g.replay()
        for k in range(len(pairs)):
            assert torch.equal(g.out[k], ref[k]), (rep, k)
        for a, b in zip(st_r, st_e):
            assert a["iterations"] == b["iterations"] and a["final_cost"] == b["final_cost"]
            assert a["solver_path_name"] == "lds"
        _say(f"[graph] replay {rep} equal")
    # new source positions in the same buffers: the replay solves them
    for k in range(len(pairs)):
        _move(g.src[k], k)
        _move(eager.src[k], k)
    st_e2 = eager.run(opt)
    st_r2 = g.replay()
    for k in range(len(pairs)):
        assert torch.equal(g.out[k], eager.out[k]), k
        assert not torch.equal(g.out[k], ref[k]), k
    assert [s["final_cost"] for s in st_r2] == [s["final_cost"] for s in st_e2]
    g.release()


def test_graph_refuses_team_pairs(gpu):
    """A pair above one workgroup needs a cooperative launch, which is not
    captured: the call reports it after its eager run, whose result is the
    ordinary one."""
    _say("[graph] case=12000-vertex pair (team) refused")
    pairs = [synth.box_union_pair(70, n_verts=12000)]
    opt = deform.default_options(max_num_iterations=3)
    b = deform.PairBatch(pairs, grid_resolution=32)
    with pytest.raises(_lib.MoError, match="cooperative"):
        b.capture(opt)
    ref = deform.PairBatch(pairs, grid_resolution=32)
    ref.run(opt)
    assert torch.equal(b.out[0], ref.out[0])


def test_graph_replay_reports_changed_inputs(gpu):
    """A target face index made out of range after the capture: the replay's
    grid build skips it and raises its error flag, and the replay reports
    that the steering values differ from the captured run."""
    _say("[graph] case=changed target faces")
    pairs = [synth.box_union_pair(71, n_boxes=4, n_verts=1200)]
    opt = deform.default_options(max_num_iterations=5)
    b = deform.PairBatch(pairs, grid_resolution=32)
    b.capture(opt)
    b.replay()
    b.tgt_f[0][0, 0] = b.tgt[0].shape[0] + 5
    with pytest.raises(_lib.MoError, match="differ from the captured run"):
        b.replay()
    b.release()


def test_graph_argument_checks(gpu):
    """out aliasing src, or timing on, is refused before anything runs."""
    pairs = [synth.box_union_pair(72, n_boxes=4, n_verts=800)]
    b = deform.PairBatch(pairs, grid_resolution=32)
    with pytest.raises(_lib.MoError, match="timing"):
        b.capture(deform.default_options(timing=1))
    b.out[0] = b.src[0]
    b._out_ptrs[0] = b.src[0].data_ptr()
    with pytest.raises(_lib.MoError, match="aliases"):
        b.capture(deform.default_options())
