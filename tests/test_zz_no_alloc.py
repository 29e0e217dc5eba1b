"""mo_deform_batch_noalloc (the batched LM with its scratch in the
library's fixed per-GPU array instead of device allocations) against
mo_deform_batch, bit for bit: two calls in a row, the second smaller, on a
batch that takes the register/LDS kernel, its multi-workgroup team form
(a 10242-vertex source) and an explicit edge list; a batch over the array's
size is refused before anything runs."""
import numpy as np
import pytest
import torch

from meshode_amd import _lib, deform, pyDeform, synth

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, flush=True)


def _pairs():
    """(param_id, normalized FP64 source, index array, kind, lambda) cases."""
    out = []
    for seed in (0, 1):
        sV, sF, tV, tF = synth.box_union_pair(seed)
        pid = deform.template_f64(tV, tF, 64)
        _, sc, pos = pyDeform.TemplateInfo(pid)
        out.append((pid, (sV - pos) / sc, sF, 0, 1.0))
    sV, sF, tV, tF = synth.displaced_sphere_pair(11, src_level=2, tgt_level=3)
    pid = deform.template_f64(tV, tF, 32)
    _, sc, pos = pyDeform.TemplateInfo(pid)
    out.append((pid, (sV - pos) / sc, sF, 0, 0.7))
    sV, sF, tV, tF = synth.displaced_sphere_pair(31, src_level=5, tgt_level=4)
    pid = deform.template_f64(tV, tF, 48)
    _, sc, pos = pyDeform.TemplateInfo(pid)
    out.append((pid, (sV - pos) / sc, sF, 0, 1.0))
    sV, sF, tV, tF = synth.box_union_pair(4, n_boxes=4, n_verts=1000)
    pid = deform.template_f64(tV, tF, 32)
    _, sc, pos = pyDeform.TemplateInfo(pid)
    E = np.concatenate([sF[:, [0, 1]], sF[:, [1, 2]]]).astype(np.int32)
    out.append((pid, (sV - pos) / sc, E, 1, 1.3))
    return out


def _solve(cases, no_alloc, **opts):
    Vs = [torch.from_numpy(np.ascontiguousarray(c[1])).cuda() for c in cases]
    st = deform.deform_batch([c[0] for c in cases], Vs, [c[2] for c in cases], [c[3] for c in cases],
                             [c[4] for c in cases], options=deform.default_options(**opts), no_alloc=no_alloc)
    return Vs, st


def _same(cases, a, b):
    (Va, sa), (Vb, sb) = a, b
    for i in range(len(cases)):
        assert torch.equal(Va[i], Vb[i]), i
        for k in ("final_cost", "iterations", "pcg_iterations", "solver_path"):
            assert sa[i][k] == sb[i][k], (i, k)


def test_no_alloc_matches_allocating_entry(gpu):
    assert _lib.load().mo_deform_batch_noalloc_max_bytes() == 64 << 20
    cases = _pairs()
    _say("[noalloc] call 1: 5 pairs (LDS kernel, team kernel, edge list)")
    first = _solve(cases, True)
    paths = {s["solver_path_name"] for s in first[1]}
    assert {"lds", "lds_team"} <= paths, paths
    _same(cases, first, _solve(cases, False))
    _say("[noalloc] call 2: 2 smaller pairs")
    small = [cases[2], cases[4]]
    second = _solve(small, True)
    _same(small, second, _solve(small, False))
    # the first call again after the smaller one: the array's stale contents do not matter
    _same(cases, _solve(cases, True), first)


def test_no_alloc_refuses_oversized_batch(gpu):
    _say("[noalloc] oversized batch and timing are refused")
    cases = _pairs()
    big = [cases[0]] * 48                   # ~1.8 MB of scratch per 5000-vertex pair
    Vs = [torch.from_numpy(np.ascontiguousarray(c[1])).cuda() for c in big]
    before = [V.clone() for V in Vs]
    with pytest.raises(_lib.MoError, match="no-allocation entry point holds at most"):
        deform.deform_batch([c[0] for c in big], Vs, [c[2] for c in big], [0] * len(big), [1.0] * len(big),
                            no_alloc=True)
    assert all(torch.equal(a, b) for a, b in zip(Vs, before))
    with pytest.raises(_lib.MoError, match="timing"):
        _solve(cases[2:3], True, timing=1)
    # still usable afterwards
    _same(cases[2:3], _solve(cases[2:3], True), _solve(cases[2:3], False))
