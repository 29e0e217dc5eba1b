"""rigid_deform CLI drop-in (rigid_deform.cc:12-63)."""
import os
import re
import subprocess

import numpy as np
import pytest

from meshode_amd import objio, synth

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CLI = os.path.join(ROOT, "meshode_amd", "bin", "rigid_deform")


def test_usage_exits_zero():
    """argc < 5 prints the usage line and exits 0 (rigid_deform.cc:13-18); no GPU needed."""
    r = subprocess.run([CLI, "a.obj", "b.obj", "c.obj"], capture_output=True, text=True)
    assert r.returncode == 0
    assert r.stdout.startswith("./rigid_deform source.obj reference.obj output.obj [GRID_RESOLUTION=64]")


def test_obj_roundtrip_rules(tmp_path):
    """objio restates Mesh::ReadOBJ / WriteOBJ (mesh.cc:14-64)."""
    p = tmp_path / "m.obj"
    p.write_text("# c\nmtllib x\nv 1 2 3\nv 4 5 6\nvn 0 0 1\nv 7 8 9\nf 1//1 2//1 3//1\nf 3/1/2 2 1\nf 1 2\n")
    V, F = objio.read_obj(str(p))
    assert V.shape == (3, 3) and np.array_equal(V[2], [7, 8, 9])
    assert F.tolist() == [[0, 1, 2], [2, 1, 0], [0, 1, 1]]      # missing token reuses the previous one
    q = tmp_path / "o.obj"
    objio.write_obj(str(q), V * 0.1234567891, F)
    assert q.read_text().splitlines()[0] == "v 0.123457 0.246914 0.37037"
    assert q.read_text().splitlines()[3] == "f 1 2 3"


@pytest.mark.gpu
def test_cli_matches_library_and_oracle(gpu, oracle, tmp_path):
    print("[cli] case=box2000_n32", flush=True)
    sV, sF, tV, tF = synth.box_union_pair(31, n_boxes=6, n_verts=2000)
    src, tgt, out = tmp_path / "s.obj", tmp_path / "t.obj", tmp_path / "o.obj"
    objio.write_obj(str(src), sV, sF)
    objio.write_obj(str(tgt), tV, tF)
    r = subprocess.run([CLI, str(src), str(tgt), str(out), "32", "5000", "1.0"], capture_output=True, text=True,
                       timeout=120)
    assert r.returncode == 0, r.stderr
    lines = r.stdout.splitlines()
    assert lines[0] == "Source:\t\tNum vertices: 2000\tNum faces: %d" % len(sF)
    assert lines[1] == "Reference:\tNum vertices: 2000\tNum faces: %d" % len(tF)
    assert lines[-1] == "Deformed"
    final = float(re.search(r"Final cost: (\S+)", r.stdout).group(1))
    # the same pair through the library (inputs parsed back from the OBJ text)
    from meshode_amd import deform
    sV2, sF2 = objio.read_obj(str(src))
    tV2, tF2 = objio.read_obj(str(tgt))
    outs, st = deform.rigid_deform_batch([(sV2, sF2, tV2, tF2)], grid_resolution=32, lambdas=1.0)
    Vo, Fo = objio.read_obj(str(out))
    assert np.array_equal(Fo, sF2)
    assert np.allclose(Vo, outs[0].cpu().numpy(), rtol=1e-5, atol=1e-6)
    assert final == pytest.approx(st[0]["final_cost"], rel=1e-5)
    # energy parity with the oracle on the same inputs
    tn, sc, pos = oracle.normalize(tV2)
    grid, _, _ = oracle.build_grid_bvh(tn, tF2, 32)
    sn = oracle.apply_transform(sV2, sc, pos)
    E = oracle.face_edges(sF2)
    _, so, _ = oracle.lm_solve(grid, 32, sn, E, 1.0)
    eg, _, _ = oracle.lm_cost(grid, 32, (outs[0].cpu().numpy() - pos) / sc, sn, E, 1.0)
    assert abs(eg - so["final_cost"]) <= 1e-3 * so["final_cost"]


@pytest.mark.gpu
def test_cli_large_source_matches_library(gpu, oracle, tmp_path):
    """A 10242-vertex source through the CLI: above the register/LDS
    kernel's reach, so its LM runs as a team of workgroups, or, when the
    runtime refuses cooperative launches in the CLI's process (as it does
    while this test process holds the GPU), on one workgroup per pair.  The
    CLI's final energy agrees within 1e-3 with the library's in-process team
    solve of the same OBJ-parsed inputs, and the oracle's cost of the CLI's
    output agrees with the energy it prints."""
    print("[cli] case=sphere_l5_n48 (large source)", flush=True)
    N = 48
    sV, sF, tV, tF = synth.displaced_sphere_pair(31, src_level=5, tgt_level=4)
    src, tgt, out = tmp_path / "s.obj", tmp_path / "t.obj", tmp_path / "o.obj"
    objio.write_obj(str(src), sV, sF)
    objio.write_obj(str(tgt), tV, tF)
    r = subprocess.run([CLI, str(src), str(tgt), str(out), str(N), "5000", "1.0"], capture_output=True, text=True,
                       timeout=120)
    assert r.returncode == 0, r.stderr
    assert r.stdout.splitlines()[-1] == "Deformed"
    final = float(re.search(r"Final cost: (\S+)", r.stdout).group(1))
    from meshode_amd import deform
    sV2, sF2 = objio.read_obj(str(src))
    tV2, tF2 = objio.read_obj(str(tgt))
    outs, st = deform.rigid_deform_batch([(sV2, sF2, tV2, tF2)], grid_resolution=N, lambdas=1.0)
    Vo, _ = objio.read_obj(str(out))
    tn, sc, pos = oracle.normalize(tV2)
    grid, _, _ = oracle.build_grid_bvh(tn, tF2, N)
    sn = oracle.apply_transform(sV2, sc, pos)
    eg, _, _ = oracle.lm_cost(grid, N, (Vo - pos) / sc, sn, oracle.face_edges(sF2), 1.0)
    print(f"[cli] large: CLI E={final:.8g} (oracle cost of its output {eg:.8g}), library team E="
          f"{st[0]['final_cost']:.8g}", flush=True)
    assert abs(final - st[0]["final_cost"]) <= 1e-3 * st[0]["final_cost"]
    assert abs(eg - final) <= 1e-4 * final


def test_load_mesh_returns_raw_coordinates(tmp_path):
    """pyDeform.LoadMesh returns the OBJ's coordinates as they are (the
    reference's src.Normalize() there is dead code, mesh_tensor.cc:95-97)
    as float32 and 0-based int32 faces; SaveMesh writes them back."""
    from meshode_amd import pyDeform
    p = tmp_path / "raw.obj"
    p.write_text("v 120.5 -3.25 7\nv 121 -3 8\nv 119 -4 7.5\nf 1 2 3\n")
    V, F = pyDeform.LoadMesh(str(p))
    assert V.dtype.is_floating_point and V.dtype.itemsize == 4 and F.dtype.itemsize == 4
    assert V.numpy().tolist() == [[120.5, -3.25, 7.0], [121.0, -3.0, 8.0], [119.0, -4.0, 7.5]]
    assert F.numpy().tolist() == [[0, 1, 2]]
    q = tmp_path / "back.obj"
    pyDeform.SaveMesh(str(q), V, F)
    V2, F2 = pyDeform.LoadMesh(str(q))
    assert V2.numpy().tolist() == V.numpy().tolist() and F2.numpy().tolist() == [[0, 1, 2]]
