"""GPU parity of the batched LM deformation (Deformer::Deform,
deformer.cc:18-92) against the oracle's FP64 LM with exact solves.

Bar: the final deformation energy of the GPU result, evaluated by the
oracle's FP64 cost function, is within 1e-3 relative of the oracle LM's
final energy on the same inputs (same FP64 grid, same normalized source)."""
import numpy as np
import pytest
import torch

from meshode_amd import deform, pyDeform, synth

pytestmark = pytest.mark.gpu

ENERGY_RTOL = 1e-3


def _say(msg):
    print(msg, flush=True)


def _setup(oracle, pair, N):
    sV, sF, tV, tF = pair
    pid = deform.template_f64(tV, tF, N)
    _, scale, pos = pyDeform.TemplateInfo(pid)
    tn, sc, ps = oracle.normalize(tV)
    assert sc == scale and np.array_equal(ps, pos)
    grid = pyDeform.TemplateGrid(pid)[0].cpu().numpy().reshape(-1)
    sn = oracle.apply_transform(sV, scale, pos)
    return pid, grid, sn, scale, pos


@pytest.mark.parametrize("case", ["sphere_l2_n32", "box1000_n32"])
def test_lm_energy_parity_small(gpu, oracle, case):
    _say(f"[lm] case={case}")
    if case == "sphere_l2_n32":
        pair = synth.displaced_sphere_pair(11, src_level=2, tgt_level=3)
    else:
        pair = synth.box_union_pair(3, n_boxes=4, n_verts=1000)
    N = 32
    pid, grid, sn, scale, pos = _setup(oracle, pair, N)
    E = oracle.face_edges(pair[1])
    xo, so, _ = oracle.lm_solve(grid, N, sn, E, 1.0)
    V = torch.from_numpy(sn.copy()).cuda()
    st = deform.deform_batch([pid], [V], [pair[1]], [0], [1.0])[0]
    xg = V.cpu().numpy()
    eg, _, _ = oracle.lm_cost(grid, N, xg, sn, E, 1.0)
    _say(f"[lm] {case}: oracle E={so['final_cost']:.8g} it={so['iterations']}  gpu E={eg:.8g} "
         f"(own {st['final_cost']:.8g}) it={st['iterations']} pcg={st['pcg_iterations']}")
    assert st["final_cost"] < st["initial_cost"]
    assert abs(eg - so["final_cost"]) <= ENERGY_RTOL * so["final_cost"]
    assert abs(st["final_cost"] - eg) <= 1e-4 * eg


def test_lm_energy_parity_headline(gpu, oracle):
    """The headline pair: 5000-vertex box union vs its jittered target, 64^3."""
    _say("[lm] case=box5000_n64")
    pair = synth.box_union_pair(0)
    N = 64
    pid, grid, sn, scale, pos = _setup(oracle, pair, N)
    E = oracle.face_edges(pair[1])
    xo, so, _ = oracle.lm_solve(grid, N, sn, E, 1.0)
    outs, st = deform.rigid_deform_batch([pair], grid_resolution=N, lambdas=1.0)
    xg = (outs[0].cpu().numpy() - pos) / scale
    eg, _, _ = oracle.lm_cost(grid, N, xg, sn, E, 1.0)
    _say(f"[lm] headline: oracle E={so['final_cost']:.8g} it={so['iterations']}  gpu E={eg:.8g} "
         f"it={st[0]['iterations']} pcg={st[0]['pcg_iterations']}")
    assert abs(eg - so["final_cost"]) <= ENERGY_RTOL * so["final_cost"]


def test_alone_equals_in_the_mix(gpu):
    """Each pair's result is bit-identical alone and inside a mixed batch."""
    _say("[lm] case=alone_vs_mix")
    pairs = [synth.box_union_pair(21, n_boxes=6, n_verts=2000),
             synth.displaced_sphere_pair(22, src_level=3, tgt_level=4),
             synth.box_union_pair(23),
             synth.displaced_sphere_pair(24, src_level=2, tgt_level=3)]
    grids = [64, 32, 64, 48]
    lams = [1.0, 0.5, 2.0, 1.0]
    mix, st_mix = deform.rigid_deform_batch(pairs, grid_resolution=grids, lambdas=lams)
    for i, p in enumerate(pairs):
        alone, st = deform.rigid_deform_batch([p], grid_resolution=grids[i], lambdas=lams[i])
        assert torch.equal(alone[0], mix[i]), i
        assert st[0]["final_cost"] == st_mix[i]["final_cost"]


def _parity(oracle, pair, N, lam=1.0, max_iterations=None, **gpu_opts):
    """Solve one pair on the GPU (deform_batch) and with the oracle's
    exact-step LM on the same grid and source; return (oracle stats, GPU
    energy under the oracle's FP64 cost, GPU stats)."""
    pid, grid, sn, scale, pos = _setup(oracle, pair, N)
    E = oracle.face_edges(pair[1])
    oo = oracle.default_lm_options()
    if max_iterations is not None:
        oo.max_num_iterations = max_iterations
        gpu_opts["max_num_iterations"] = max_iterations
    xo, so, _ = oracle.lm_solve(grid, N, sn, E, lam, options=oo)
    V = torch.from_numpy(sn.copy()).cuda()
    st = deform.deform_batch([pid], [V], [pair[1]], [0], [lam], options=deform.default_options(**gpu_opts))[0]
    eg, _, _ = oracle.lm_cost(grid, N, V.cpu().numpy(), sn, E, lam)
    return so, eg, st


@pytest.mark.parametrize("case", ["sphere_l2_n32", "box1000_n32"])
def test_lm_global_kernel_energy_parity(gpu, oracle, case):
    """The global-memory kernel (k_lm_solve, kernel_variant 1) on the small
    cases: same energy bar as the register/LDS kernel."""
    _say(f"[lm] case={case} kernel_variant=1")
    if case == "sphere_l2_n32":
        pair = synth.displaced_sphere_pair(11, src_level=2, tgt_level=3)
    else:
        pair = synth.box_union_pair(3, n_boxes=4, n_verts=1000)
    so, eg, st = _parity(oracle, pair, 32, kernel_variant=1)
    _say(f"[lm] {case} global: oracle E={so['final_cost']:.8g}  gpu E={eg:.8g} it={st['iterations']}")
    assert st["final_cost"] < st["initial_cost"]
    assert abs(eg - so["final_cost"]) <= ENERGY_RTOL * so["final_cost"]
    assert abs(st["final_cost"] - eg) <= 1e-4 * eg


def test_lm_nonmanifold_edge_energy_parity(gpu, oracle):
    """A fin triangle on an existing edge gives that vertex pair three
    directed edges (the register/LDS kernel stores multiplicities 1 and 2
    only), so the pair goes to the global-memory kernel on its own."""
    _say("[lm] case=fin_l2_n32")
    sV, sF, tV, tF = synth.displaced_sphere_pair(12, src_level=2, tgt_level=3)
    a, b = int(sF[0, 0]), int(sF[0, 1])
    mid = 0.5 * (sV[a] + sV[b])
    fin = mid + 0.3 * mid / np.linalg.norm(mid)
    sV2 = np.vstack([sV, fin[None]])
    sF2 = np.vstack([sF, np.array([[a, b, len(sV)]], sF.dtype)])
    E = oracle.face_edges(sF2)
    und = np.sort(E, 1)
    _, counts = np.unique(und, axis=0, return_counts=True)
    assert counts.max() == 3
    so, eg, st = _parity(oracle, (sV2, sF2, tV, tF), 32)
    _say(f"[lm] fin: oracle E={so['final_cost']:.8g}  gpu E={eg:.8g} it={st['iterations']}")
    assert st["final_cost"] < st["initial_cost"]
    assert abs(eg - so["final_cost"]) <= ENERGY_RTOL * so["final_cost"]


def test_lm_large_pair_energy_parity(gpu, oracle):
    """A 10242-vertex source (icosphere level 5, above the register/LDS
    kernel's 8192, so the global-memory kernel solves it) against a level-4
    target at 48^3.  The oracle's converged energy for this seeded pair is
    the golden fixture lm_sphere_l5_n48.npz (tests/golden/make_golden.py);
    the GPU result is evaluated with the oracle's FP64 cost on the GPU's
    template grid, which is bit-exact to the oracle's."""
    import os
    z = np.load(os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "lm_sphere_l5_n48.npz"))
    N = int(z["N"])
    _say(f"[lm] case=sphere_l5_n{N} (global kernel)")
    pair = synth.displaced_sphere_pair(int(z["seed"]), src_level=int(z["src_level"]), tgt_level=int(z["tgt_level"]))
    assert len(pair[0]) > 8192
    pid, grid, sn, scale, pos = _setup(oracle, pair, N)
    E = oracle.face_edges(pair[1])
    V = torch.from_numpy(sn.copy()).cuda()
    st = deform.deform_batch([pid], [V], [pair[1]], [0], [float(z["lam"])])[0]
    eg, _, _ = oracle.lm_cost(grid, N, V.cpu().numpy(), sn, E, float(z["lam"]))
    ref = float(z["final_cost"])
    _say(f"[lm] large: oracle E={ref:.8g} it={int(z['iterations'])}  gpu E={eg:.8g} it={st['iterations']}")
    assert abs(eg - ref) <= ENERGY_RTOL * ref
    assert abs(st["final_cost"] - eg) <= 1e-4 * eg


@pytest.mark.parametrize("case", ["sphere_l2_n32", "box1000_n32", "ico4_n32"])
def test_lm_team_kernel_energy_parity(gpu, oracle, case):
    """The multi-workgroup kernel (k_lm_solve_team) forced with 256-thread
    workgroups (kernel_variant 2: a team of ceil(nv / 256)): a team of one
    for the 162-vertex sphere, 4 workgroups for the box union, 11 for the
    2562-vertex icosphere.  Same energy bar as the other kernels."""
    _say(f"[lm] case={case} kernel_variant=2 block_threads=256")
    if case == "sphere_l2_n32":
        pair = synth.displaced_sphere_pair(11, src_level=2, tgt_level=3)
    elif case == "box1000_n32":
        pair = synth.box_union_pair(3, n_boxes=4, n_verts=1000)
    else:
        pair = synth.displaced_sphere_pair(13, src_level=4, tgt_level=4)
    so, eg, st = _parity(oracle, pair, 32, kernel_variant=2, block_threads=256)
    _say(f"[lm] {case} team: oracle E={so['final_cost']:.8g}  gpu E={eg:.8g} it={st['iterations']}")
    assert st["final_cost"] < st["initial_cost"]
    assert abs(eg - so["final_cost"]) <= ENERGY_RTOL * so["final_cost"]
    assert abs(st["final_cost"] - eg) <= 1e-4 * eg


def test_team_kernel_alone_equals_in_the_mix(gpu):
    """A team's size depends on its pair alone, so each result is
    bit-identical alone, inside a mixed batch, and on a second run."""
    _say("[lm] case=team alone_vs_mix")
    pairs = [synth.box_union_pair(21, n_boxes=6, n_verts=2000),
             synth.displaced_sphere_pair(22, src_level=3, tgt_level=4),
             synth.displaced_sphere_pair(24, src_level=2, tgt_level=3)]
    opt = deform.default_options(kernel_variant=2, block_threads=256)
    mix, st_mix = deform.rigid_deform_batch(pairs, grid_resolution=[48, 32, 32], lambdas=[1.0, 0.5, 2.0],
                                            options=opt)
    again, _ = deform.rigid_deform_batch(pairs, grid_resolution=[48, 32, 32], lambdas=[1.0, 0.5, 2.0], options=opt)
    for i, p in enumerate(pairs):
        alone, st = deform.rigid_deform_batch([p], grid_resolution=[48, 32, 32][i], lambdas=[1.0, 0.5, 2.0][i],
                                              options=opt)
        assert torch.equal(alone[0], mix[i]), i
        assert torch.equal(again[i], mix[i]), i
        assert st[0]["final_cost"] == st_mix[i]["final_cost"]


def test_graph_kind_on_team_kernel(gpu, oracle):
    """DeformGraph's input (kind 1: an explicit directed edge list) on the
    multi-workgroup kernel (kernel_variant 2, 128-thread workgroups: a team
    of 8 for the 1000-vertex box union), against the oracle's LM on the same
    edges."""
    _say("[lm] case=graph kind 1 on a team")
    sV, sF, tV, tF = synth.box_union_pair(4, n_boxes=4, n_verts=1000)
    N = 32
    pid, grid, sn, scale, pos = _setup(oracle, (sV, sF, tV, tF), N)
    E = oracle.face_edges(sF)
    _, so, _ = oracle.lm_solve(grid, N, sn, E, 0.7)
    V = torch.from_numpy(sn.copy()).cuda()
    st = deform.deform_batch([pid], [V], [torch.from_numpy(E).cuda()], [1], [0.7],
                             options=deform.default_options(kernel_variant=2, block_threads=128))[0]
    eg, _, _ = oracle.lm_cost(grid, N, V.cpu().numpy(), sn, E, 0.7)
    _say(f"[lm] graph team: oracle E={so['final_cost']:.8g}  gpu E={eg:.8g} it={st['iterations']}")
    assert abs(eg - so["final_cost"]) <= ENERGY_RTOL * so["final_cost"]


def test_team_barrier_timeout_is_reported(gpu, monkeypatch):
    """A team barrier wait that exceeds its bound (forced from the host:
    MO_LM_TEAM_WAIT_US=-1 makes every wait give up at once) stops all
    waiting, the kernel still finishes, and the call raises; the next call
    with the normal bound solves as before."""
    from meshode_amd import _lib
    _say("[lm] case=team barrier bound forced")
    pair = synth.box_union_pair(3, n_boxes=4, n_verts=1000)
    opt = deform.default_options(kernel_variant=2, block_threads=64, max_num_iterations=5)
    ref, _ = deform.rigid_deform_batch([pair], grid_resolution=32, options=opt)
    monkeypatch.setenv("MO_LM_TEAM_WAIT_US", "-1")
    with pytest.raises(_lib.MoError, match="barrier wait"):
        deform.rigid_deform_batch([pair], grid_resolution=32, options=opt)
    # the in-place entry point leaves the caller's vertices untouched
    pid = deform.template_f64(pair[2], pair[3], 32)
    _, scale, pos = pyDeform.TemplateInfo(pid)
    V = torch.from_numpy((pair[0] - pos) / scale).cuda()
    V_before = V.clone()
    with pytest.raises(_lib.MoError, match="outputs not written"):
        deform.deform_batch([pid], [V], [pair[1]], [0], [1.0], options=opt)
    assert torch.equal(V, V_before)
    monkeypatch.delenv("MO_LM_TEAM_WAIT_US")
    again, _ = deform.rigid_deform_batch([pair], grid_resolution=32, options=opt)
    assert torch.equal(ref[0], again[0])


def test_team_with_idle_workgroups(gpu, oracle):
    """team_vertices 1 gives the 162-vertex sphere a team of 32 workgroups of
    64 threads, of which only the first 3 own vertices; the idle members
    still take part in every reduction and barrier."""
    _say("[lm] case=team of 32 with 29 idle workgroups")
    pair = synth.displaced_sphere_pair(11, src_level=2, tgt_level=3)
    so, eg, st = _parity(oracle, pair, 32, kernel_variant=2, block_threads=64, team_vertices=1)
    _say(f"[lm] idle team: oracle E={so['final_cost']:.8g}  gpu E={eg:.8g} it={st['iterations']}")
    assert abs(eg - so["final_cost"]) <= ENERGY_RTOL * so["final_cost"]


def test_solver_path_is_reported(gpu, monkeypatch):
    """mo_lm_stats.solver_path names the kernel that solved each pair; a
    refused cooperative launch (forced from the host with
    MO_LM_COOP_REFUSE=1) is reported as such and gives the one-workgroup
    kernel's result bit for bit."""
    _say("[lm] case=solver path reporting")
    small = synth.box_union_pair(3, n_boxes=4, n_verts=1000)
    _, st = deform.rigid_deform_batch([small], grid_resolution=32)
    assert st[0]["solver_path_name"] == "lds"
    one, st1 = deform.rigid_deform_batch([small], grid_resolution=32, options=deform.default_options(kernel_variant=1))
    assert st1[0]["solver_path_name"] == "one_workgroup"
    team_opt = deform.default_options(kernel_variant=2, block_threads=128)
    _, st2 = deform.rigid_deform_batch([small], grid_resolution=32, options=team_opt)
    assert st2[0]["solver_path_name"] == "team"
    monkeypatch.setenv("MO_LM_COOP_REFUSE", "1")
    fb, st3 = deform.rigid_deform_batch([small], grid_resolution=32, options=team_opt)
    monkeypatch.delenv("MO_LM_COOP_REFUSE")
    assert st3[0]["solver_path_name"] == "one_workgroup_coop_refused"
    ref1, _ = deform.rigid_deform_batch([small], grid_resolution=32,
                                        options=deform.default_options(kernel_variant=1, block_threads=128))
    assert torch.equal(fb[0], ref1[0])


def test_dense_graph_beyond_lds_offsets(gpu, oracle):
    """A complete graph on 270 vertices (kind 1) has 72,630 adjacency
    entries: it fits the LDS budget, but its LDS offsets do not fit the
    register/LDS kernel's 16-bit packing, so it must go to the global
    kernel, and still meet the energy bar."""
    _say("[lm] case=K270 dense graph")
    rng = np.random.default_rng(31)
    _, _, tV, tF = synth.displaced_sphere_pair(12, src_level=2, tgt_level=3)
    N = 32
    pid = deform.template_f64(tV, tF, N)
    grid = pyDeform.TemplateGrid(pid)[0].cpu().numpy().reshape(-1)
    n = 270
    sn = rng.uniform(0.2, 0.8, (n, 3))
    ii, jj = np.triu_indices(n, 1)
    E = np.stack([ii, jj], 1).astype(np.int32)
    lam = 0.05
    _, so, _ = oracle.lm_solve(grid, N, sn, E, lam)
    V = torch.from_numpy(sn.copy()).cuda()
    st = deform.deform_batch([pid], [V], [torch.from_numpy(E).cuda()], [1], [lam])[0]
    eg, _, _ = oracle.lm_cost(grid, N, V.cpu().numpy(), sn, E, lam)
    _say(f"[lm] K270: oracle E={so['final_cost']:.8g}  gpu E={eg:.8g} it={st['iterations']} "
         f"path={st['solver_path_name']}")
    assert st["solver_path_name"] == "one_workgroup"
    assert abs(eg - so["final_cost"]) <= ENERGY_RTOL * so["final_cost"]


def test_lds_team_barrier_timeout_is_reported(gpu, monkeypatch):
    """The register/LDS team kernel (a 12000-vertex pair: a team of 4
    workgroups of 3000 vertices) with every barrier wait forced to give up at once
    (MO_LM_TEAM_WAIT_US=-1, from the host): the kernel finishes, the call
    raises and leaves the caller's vertices untouched; the next call solves
    bit-identically to one made before."""
    from meshode_amd import _lib
    _say("[lm] case=register/LDS team barrier bound forced")
    pair = synth.box_union_pair(5, n_verts=12000)
    pid = deform.template_f64(pair[2], pair[3], 32)
    _, scale, pos = pyDeform.TemplateInfo(pid)
    x0 = (pair[0] - pos) / scale
    opt = deform.default_options(max_num_iterations=5)
    V = torch.from_numpy(x0.copy()).cuda()
    st = deform.deform_batch([pid], [V], [pair[1]], [0], [1.0], options=opt)[0]
    assert st["solver_path_name"] == "lds_team"
    monkeypatch.setenv("MO_LM_TEAM_WAIT_US", "-1")
    W = torch.from_numpy(x0.copy()).cuda()
    with pytest.raises(_lib.MoError, match="outputs not written"):
        deform.deform_batch([pid], [W], [pair[1]], [0], [1.0], options=opt)
    assert torch.equal(W, torch.from_numpy(x0).cuda())
    monkeypatch.delenv("MO_LM_TEAM_WAIT_US")
    deform.deform_batch([pid], [W], [pair[1]], [0], [1.0], options=opt)
    assert torch.equal(W, V)


def test_lds_team_kernel_energy_parity(gpu, oracle):
    """A 12000-vertex box union at 48^3 on the register/LDS team kernel
    (4 workgroups of 3000 vertices, neighbours across chunks read from the
    published array) against the oracle's exact-step LM."""
    _say("[lm] case=box12000_n48 register/LDS team")
    pair = synth.box_union_pair(6, n_verts=12000)
    so, eg, st = _parity(oracle, pair, 48)
    _say(f"[lm] lds team: oracle E={so['final_cost']:.8g}  gpu E={eg:.8g} it={st['iterations']} "
         f"path={st['solver_path_name']}")
    assert st["solver_path_name"] == "lds_team" and st["workgroups"] == 4
    assert abs(eg - so["final_cost"]) <= ENERGY_RTOL * so["final_cost"]


def test_lds_team_many_teams_deterministic(gpu):
    """Eight pairs of 9000-30000 vertices (teams of 3 to 8 workgroups in one
    cooperative grid) solved twice: every pair's vertices are bit-identical
    between the calls, and each pair matches its own solve alone (the
    teams' exchanges through the publication array and the partial slots
    carry no stale value)."""
    _say("[lm] case=8 register/LDS teams, repeat and alone")
    pairs = [synth.box_union_pair(40 + k, n_verts=9000 + 3000 * k) for k in range(8)]
    opt = deform.default_options(max_num_iterations=8)
    a, sa = deform.rigid_deform_batch(pairs, grid_resolution=32, options=opt)
    b, _ = deform.rigid_deform_batch(pairs, grid_resolution=32, options=opt)
    assert all(s["solver_path_name"] == "lds_team" for s in sa), [s["solver_path_name"] for s in sa]
    for k in range(8):
        assert torch.equal(a[k], b[k]), k
    for k in (0, 7):
        alone, _ = deform.rigid_deform_batch([pairs[k]], grid_resolution=32, options=opt)
        assert torch.equal(alone[0], a[k]), k
