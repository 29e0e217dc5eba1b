"""CAD chain on the GPU: LinearEstimation (linear.cc:11-112) against the
oracle's sparse direct solve of the same system, pyDeform.SolveLinear and
LoadCadMesh, and the cad_deform pipeline (cad_deform.cc:30-120): the
representative-graph solve against the oracle LM, the propagation against
the oracle's LinearEstimation, and the CLI against the in-process run."""
import os
import subprocess

import numpy as np
import pytest
import torch

from meshode_amd import cad, objio, synth

from oracle import cad_oracle

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _cad_box(size=(0.3, 0.2, 0.25), angle=0.4):
    x, y, z = size
    V = np.array([[0, 0, 0], [x, 0, 0], [x, y, 0], [0, y, 0], [0, 0, z], [x, 0, z], [x, y, z], [0, y, z]], float)
    F = np.array([[0, 2, 1], [0, 3, 2], [4, 5, 6], [4, 6, 7], [0, 1, 5], [0, 5, 4], [1, 2, 6], [1, 6, 5],
                  [2, 3, 7], [2, 7, 6], [3, 0, 4], [3, 4, 7]], np.int32)
    c, s = np.cos(angle), np.sin(angle)
    R = np.array([[c, -s, 0], [s, c, 0], [0, 0, 1]])
    return V @ R.T + 0.05, F


@pytest.mark.gpu
def test_linear_estimation_against_direct_solve(gpu):
    print("[cad] case=linear_estimation rotated box", flush=True)
    V, F = _cad_box()
    r = cad.prepare(V, F)
    rng = np.random.default_rng(3)
    GV = r["GV"] + rng.normal(0, 2e-3, r["GV"].shape)
    want = cad_oracle.linear_estimation(r["V"], r["F"], r["E"], r["ref"], GV, 2.0)
    Vd = torch.from_numpy(r["V"]).cuda()
    it = cad.linear_estimation(Vd, torch.from_numpy(r["F"]).cuda(), torch.from_numpy(r["E"]).cuda(),
                               torch.from_numpy(r["ref"]).cuda(), torch.from_numpy(GV).cuda(), 2.0)
    got = Vd.cpu().numpy()
    assert it > 0
    np.testing.assert_allclose(got, want, rtol=0, atol=1e-9)


@pytest.mark.gpu
def test_solve_linear_and_load_cad_mesh(gpu, tmp_path):
    from meshode_amd import pyDeform
    print("[cad] case=pyDeform LoadCadMesh + SolveLinear", flush=True)
    V, F = _cad_box((0.2, 0.15, 0.1), 0.2)
    path = str(tmp_path / "cad.obj")
    objio.write_obj(path, V, F)
    Vr, Fr = objio.read_obj(path)
    out = pyDeform.LoadCadMesh(path)
    r = cad.prepare(Vr, Fr)
    assert [t.dtype for t in out] == [torch.float32, torch.int32, torch.int32, torch.int32, torch.float32,
                                      torch.int32]
    np.testing.assert_array_equal(out[0].numpy(), r["V"].astype(np.float32))
    np.testing.assert_array_equal(out[3].numpy().ravel(), r["ref"])
    np.testing.assert_array_equal(out[5].numpy(), r["GE"])
    src_V, src_F, src_E, s2g, gV = out[0].clone(), out[1], out[2], out[3], out[4] + 1e-3
    pyDeform.SolveLinear(src_V, src_F, src_E, s2g, gV, 1.0)
    want = cad_oracle.linear_estimation(out[0].double().numpy(), src_F.numpy(), src_E.numpy(), s2g.numpy().ravel(),
                                        gV.double().numpy(), 1.0)
    np.testing.assert_allclose(src_V.numpy(), want.astype(np.float32), rtol=0, atol=2e-7)


def _target():
    sV, sF, tV, tF = synth.displaced_sphere_pair(7, src_level=2, tgt_level=3)
    return tV * 0.15 + 0.1, tF


@pytest.mark.gpu
def test_cad_deform_pipeline_against_oracle(gpu, oracle):
    print("[cad] case=cad_deform pipeline box -> sphere", flush=True)
    V, F = _cad_box()
    tV, tF = _target()
    N = 32
    res = cad.cad_deform(V, F, tV, tF, grid_resolution=N, lam=1.0)
    try:
        from meshode_amd import pyDeform
        grid = pyDeform.TemplateGrid(res["pid"])[0].cpu().numpy().reshape(-1)
        # DeformGraph energy against the oracle LM on the same graph and grid
        GE = res["prep"]["GE"]
        xo, st, _ = oracle.lm_solve(grid, N, res["GV0"], GE, 1.0)
        eg, _, _ = oracle.lm_cost(grid, N, res["GV"], res["GV0"], GE, 1.0)
        eo, _, _ = oracle.lm_cost(grid, N, xo, res["GV0"], GE, 1.0)
        assert eg == pytest.approx(eo, rel=1e-3)
        assert res["stats"]["final_cost"] == pytest.approx(eg, rel=1e-5)
        # propagation: the oracle's LinearEstimation from the GPU's deformed graph
        info = pyDeform.TemplateInfo(res["pid"])
        scale, pos = info[1], np.asarray(info[2])
        Vn = (res["prep"]["V"] - pos) / scale
        want = cad_oracle.linear_estimation(Vn, res["prep"]["F"], res["prep"]["E"], res["prep"]["ref"], res["GV"],
                                            2.0)
        np.testing.assert_allclose(res["V"], want * scale + pos, rtol=0, atol=1e-8)
    finally:
        from meshode_amd import _lib
        _lib.call("mo_template_destroy", res["pid"])


@pytest.mark.gpu
def test_cad_deform_cli(gpu, tmp_path):
    exe = os.path.join(ROOT, "meshode_amd", "bin", "cad_deform")
    if not os.path.exists(exe):
        pytest.skip("cad_deform not built")
    print("[cad] case=cad_deform CLI", flush=True)
    V, F = _cad_box()
    tV, tF = _target()
    src, tgt, out = str(tmp_path / "cad.obj"), str(tmp_path / "ref.obj"), str(tmp_path / "out.obj")
    objio.write_obj(src, V, F)
    objio.write_obj(tgt, tV, tF)
    p = subprocess.run([exe, src, tgt, out, "32", "5000", "1.0", "0"], capture_output=True, text=True, timeout=300)
    assert p.returncode == 0, p.stderr
    lines = p.stdout.splitlines()
    assert lines[0].startswith("Source:\t\tNum vertices: 8\tNum faces: 12")
    assert any(l.startswith("Vertices cost: ") for l in lines) and lines[-1] == "Deformed"
    Vo, Fo = objio.read_obj(out)
    Vr, Fr = objio.read_obj(src)
    tr, tf = objio.read_obj(tgt)
    res = cad.cad_deform(Vr, Fr, tr, tf, grid_resolution=32, lam=1.0)
    from meshode_amd import _lib
    _lib.call("mo_template_destroy", res["pid"])
    np.testing.assert_array_equal(Fo, res["F"])
    # the OBJ text keeps 6 significant digits (std::ostream default)
    np.testing.assert_allclose(Vo, res["V"], rtol=2e-6, atol=1e-6)
    # usage line and exit 0 with too few arguments
    p = subprocess.run([exe, src], capture_output=True, text=True, timeout=60)
    assert p.returncode == 0 and p.stdout.startswith("./cad_deform cad.obj reference.obj output.obj")


@pytest.mark.gpu
@pytest.mark.parametrize("team", [False, True])
def test_adaptive_edge_lm_against_oracle(gpu, oracle, team):
    """kind 2 (AdaptiveEdgeLoss weights) through mo_deform_batch: full solve
    and a solve capped at two successful steps, against the oracle; with
    `team`, on the multi-workgroup kernel (kernel_variant 2, 64-thread
    workgroups: a team of 3 for the 162 vertices)."""
    from meshode_amd import deform, pyDeform
    print(f"[cad] case=adaptive-weight LM sphere team={team}", flush=True)
    kopt = {"kernel_variant": 2, "block_threads": 64} if team else {}
    sV, sF, tV, tF = synth.displaced_sphere_pair(5, src_level=2, tgt_level=3)
    N = 32
    pid = deform.template_f64(tV, tF, N)
    try:
        _, scale, pos = pyDeform.TemplateInfo(pid)
        grid = pyDeform.TemplateGrid(pid)[0].cpu().numpy().reshape(-1)
        sn = (sV - pos) / scale
        E = oracle.face_edges(sF)
        lam = oracle.adaptive_lambda(sn, E, 1.0)
        for cap in (0, 2):
            V = torch.from_numpy(sn.copy()).cuda()
            st = deform.deform_batch([pid], [V], [torch.from_numpy(E).cuda()], [2], [1.0],
                                     options=deform.default_options(max_successful_steps=cap, **kopt))[0]
            xo, so = oracle.lm_solve_ex(grid, N, sn, sn, E, lam_e=lam, max_successful=cap)
            eg, _, _ = oracle.lm_cost_ex(grid, N, V.cpu().numpy(), sn, E, lam)
            assert st["final_cost"] == pytest.approx(eg, rel=1e-5)
            assert eg == pytest.approx(so["final_cost"], rel=1e-3 if cap == 0 else 5e-3)
            if cap:
                assert st["successful_steps"] == cap == so["successful_steps"]
    finally:
        from meshode_amd import _lib
        _lib.call("mo_template_destroy", pid)


def _oracle_subdivision_loop(oracle, grid, N, Vn, E, lam_e):
    """The repeated-solve loop of Deformer::DeformSubdivision with the
    TerminateWhenSuccessCallback (deformer.cc:229-243, callback.h:22-36)."""
    counter, prev, snaps, x = 0, 1e30, [], Vn.copy()
    for _ in range(100000):
        xn, st = oracle.lm_solve_ex(grid, N, x, Vn, E, lam_e=lam_e, max_successful=2 if counter == 0 else 1)
        k = st["successful_steps"]
        if counter == 0:
            if k >= 1:
                snaps.append(x.copy())
            counter = 1 if k == 1 else 0
        else:
            counter = 0 if k >= 1 else 1
        x = xn
        if abs(prev - st["final_cost"]) < 1e-6:
            break
        prev = st["final_cost"]
    return x, st, snaps


@pytest.mark.gpu
def test_deform_subdivision_flow_against_oracle(gpu, oracle):
    from meshode_amd import deform, pyDeform
    print("[cad] case=DeformSubdivision flow box -> sphere", flush=True)
    V, F = _cad_box((0.2, 0.15, 0.1), 0.3)
    tV, tF = _target()
    r = cad.prepare(V, F)
    N = 32
    pid = deform.template_f64(tV, tF, N)
    try:
        _, scale, pos = pyDeform.TemplateInfo(pid)
        grid = pyDeform.TemplateGrid(pid)[0].cpu().numpy().reshape(-1)
        Vn = (r["V"] - pos) / scale
        Eall = np.concatenate([r["E"], np.stack([r["F"][:, [0, 1, 2]].ravel(), r["F"][:, [1, 2, 0]].ravel()], 1)])
        lam = oracle.adaptive_lambda(Vn, Eall, 1.0)
        rest = torch.from_numpy(Vn).cuda()
        Vd = rest.clone()
        st, snaps = cad.deform_subdivision(pid, rest, Vd, torch.from_numpy(r["F"]).cuda(),
                                           torch.from_numpy(r["E"]).cuda(), 1.0)
        xo, so, osnaps = _oracle_subdivision_loop(oracle, grid, N, Vn, Eall, lam)
        assert len(snaps) >= 1 and np.array_equal(snaps[0], Vn)     # the first callback sees the start
        eg, _, _ = oracle.lm_cost_ex(grid, N, Vd.cpu().numpy(), Vn, Eall, lam)
        assert st["final_cost"] == pytest.approx(eg, rel=1e-5)
        assert eg == pytest.approx(so["final_cost"], rel=1e-3)
        print("[cad] solves", st["solves"], "snapshots gpu/oracle", len(snaps), len(osnaps), flush=True)
    finally:
        from meshode_amd import _lib
        _lib.call("mo_template_destroy", pid)


@pytest.mark.gpu
def test_cad_deform_cli_flow_output(gpu, tmp_path):
    exe = os.path.join(ROOT, "meshode_amd", "bin", "cad_deform")
    if not os.path.exists(exe):
        pytest.skip("cad_deform not built")
    print("[cad] case=cad_deform CLI flow", flush=True)
    V, F = _cad_box((0.2, 0.15, 0.1), 0.3)
    tV, tF = _target()
    src, tgt = str(tmp_path / "cad.obj"), str(tmp_path / "ref.obj")
    out, flow = str(tmp_path / "out.obj"), str(tmp_path / "flow.obj")
    objio.write_obj(src, V, F)
    objio.write_obj(tgt, tV, tF)
    p = subprocess.run([exe, src, tgt, out, "32", "5000", "1.0", "0", flow], capture_output=True, text=True,
                       timeout=600)
    assert p.returncode == 0, p.stderr
    assert p.stdout.splitlines()[-1] == "Deformed"
    Vo, Fo = objio.read_obj(out)
    lines = open(flow).read().splitlines()
    nv = len(Vo)
    nvl = sum(1 for l in lines if l.startswith("v "))
    nll = sum(1 for l in lines if l.startswith("l "))
    assert nvl % nv == 0
    ns = nvl // nv
    assert ns >= 2 and nll == nv * (ns - 1)
    # the last recorded position of each vertex is the output mesh (6 digits)
    P = np.array([[float(x) for x in l.split()[1:]] for l in lines if l.startswith("v ")]).reshape(nv, ns, 3)
    np.testing.assert_allclose(P[:, -1], Vo, rtol=2e-6, atol=1e-6)
