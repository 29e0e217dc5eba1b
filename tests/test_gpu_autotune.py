"""GPU: every launch shape of the register/LDS LM kernel meets the energy bar
against the oracle (final energy within 1e-3 relative, as for the default
shape), and the autotuner installs a size-class table whose choice is what
later calls run (bit-identical to forcing that shape)."""
import numpy as np
import pytest
import torch

from meshode_amd import autotune, deform, pyDeform, synth

pytestmark = pytest.mark.gpu

ENERGY_RTOL = 1e-3


def _say(msg):
    print(msg, flush=True)


def test_every_lds_shape_meets_the_energy_bar(gpu, oracle):
    """A 1000-vertex box union fits all 13 shapes (T * VPT >= 1024)."""
    _say("[autotune] case=box1000_n32 every shape")
    sV, sF, tV, tF = synth.box_union_pair(5, n_boxes=4, n_verts=1000)
    N = 32
    pid = deform.template_f64(tV, tF, N)
    _, scale, pos = pyDeform.TemplateInfo(pid)
    grid = pyDeform.TemplateGrid(pid)[0].cpu().numpy().reshape(-1)
    sn = oracle.apply_transform(sV, scale, pos)
    E = oracle.face_edges(sF)
    _, so, _ = oracle.lm_solve(grid, N, sn, E, 1.0)
    try:
        for b, (t, v) in enumerate(autotune.shapes()):
            autotune.set_table([(1024, b)])
            V = torch.from_numpy(sn.copy()).cuda()
            st = deform.deform_batch([pid], [V], [sF], [0], [1.0])[0]
            eg, _, _ = oracle.lm_cost(grid, N, V.cpu().numpy(), sn, E, 1.0)
            _say(f"[autotune] {t}x{v}: gpu E={eg:.8g} oracle E={so['final_cost']:.8g} it={st['iterations']}")
            assert abs(eg - so["final_cost"]) <= ENERGY_RTOL * so["final_cost"], (t, v)
    finally:
        autotune.set_table([])


def test_tuner_installs_the_fastest_fitting_shapes(gpu):
    """Two size classes (642 and 2562 vertices): the installed table has one
    fitting row per class, and a pair solved after tuning equals the same
    pair solved with its class's shape forced."""
    _say("[autotune] case=tune L3 + L4")
    pairs = [synth.icosphere_pair(s, 3) for s in range(4)] + [synth.icosphere_pair(10 + s, 4) for s in range(4)]
    try:
        res = autotune.tune(pairs, grid_resolution=32, repeats=1)
        _say(f"[autotune] {autotune.describe(res)}")
        assert [r[0] for r in res["table"]] == [1024, 4096]
        for ub, t, v, _ in res["table"]:
            assert t * v >= ub
        assert res["block_threads"] is None
        tuned, _ = deform.rigid_deform_batch(pairs[4:5], grid_resolution=32)
        shp = autotune.shapes()
        b = shp.index((res["table"][1][1], res["table"][1][2]))
        autotune.set_table([(4096, b)])
        forced, _ = deform.rigid_deform_batch(pairs[4:5], grid_resolution=32)
        assert torch.equal(tuned[0], forced[0])
    finally:
        autotune.set_table([])


def test_tuner_block_threads_for_large_pairs(gpu):
    """A pair above 8192 vertices is timed on the multi-workgroup kernel at
    each block size and team size; the table stays empty."""
    _say("[autotune] case=tune L5 (global kernel)")
    pairs = [synth.icosphere_pair(30, 5)]
    opt = deform.default_options(max_num_iterations=3)
    try:
        res = autotune.tune(pairs, grid_resolution=32, options=opt, repeats=1, block_threads=(512, 1024),
                            team_factors=(2, 4))
        _say(f"[autotune] {autotune.describe(res)} {res['large_ms']}")
        assert res["table"] == [] and res["block_threads"] in (512, 1024)
        assert res["team_vertices"] in (2 * res["block_threads"], 4 * res["block_threads"])
    finally:
        autotune.set_table([])
