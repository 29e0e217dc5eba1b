"""Writes the small golden fixtures under tests/golden/ from the CPU oracle.

The reference ships no golden vectors and cannot be built here, so these
fixtures are the oracle's outputs on seeded inputs (the oracle itself is
pinned by closed forms, invariants and independent brute force in
tests/test_oracle_*.py).  They freeze the oracle's behaviour and give the GPU
tests fixed expected outputs.  Run: python tests/golden/make_golden.py
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

from meshode_amd import synth  # noqa: E402
from oracle import oracle as O  # noqa: E402


def main():
    # grid: displaced sphere target (level 2), float32 input as pyDeform gives it
    _, _, tV, tF = synth.displaced_sphere_pair(40, src_level=1, tgt_level=2)
    tV32 = tV.astype(np.float32)
    tn, sc, pos = O.normalize(tV32.astype(np.float64))
    N = 12
    d64, d32, nn = O.build_grid(tn, tF, N)
    np.savez_compressed(os.path.join(HERE, "grid_sphere_n12.npz"), V=tV32, F=tF, N=N, scale=sc, pos=pos,
                        dist64=d64, nearest=nn)
    # lookups / distance loss on that grid at probe points covering every branch
    rng = np.random.default_rng(41)
    P = np.concatenate([rng.uniform(0, 1, (300, 3)), rng.integers(0, N + 1, (60, 3)) / N,
                        rng.uniform(-0.3, 1.3, (100, 3)), rng.uniform((N - 1) / N, 1, (40, 3))]).astype(np.float32)
    f32d, f32g = O.lookup_f32(d64, N, P)
    np.savez_compressed(os.path.join(HERE, "lookup_n12.npz"), P=P, d=f32d, g=f32g,
                        fwd=O.distance_loss_fwd(d64, N, P), bwd=O.distance_loss_bwd(d64, N, P))
    # edge losses
    sV, sF = synth.icosphere(2)
    V0 = (sV * 0.3 + 0.5).astype(np.float32)
    V1 = (V0 + rng.normal(0, 0.01, V0.shape)).astype(np.float32)
    off = O.store_rigidity(V0, sF)
    E = rng.integers(0, len(V0), (500, 2)).astype(np.int32)
    goff = O.store_graph(V0, E)
    np.savez_compressed(os.path.join(HERE, "edges_ico2.npz"), V0=V0, V1=V1, F=sF, E=E,
                        rigid_fwd=O.rigid_fwd(V1, sF, off), rigid_bwd=O.rigid_bwd(V1, sF, off),
                        graph_fwd=O.graph_fwd(V1, E, goff), graph_bwd=O.graph_bwd(V1, E, goff))
    # one small LM solve (Deformer::Deform) on an FP64 target template
    pV, pF, qV, qF = synth.displaced_sphere_pair(42, src_level=2, tgt_level=3)
    qn, qs, qp = O.normalize(qV)
    M = 24
    g, _, _ = O.build_grid(qn, qF, M)
    x0 = O.apply_transform(pV, qs, qp)
    x, st, _ = O.lm_solve(g, M, x0, O.face_edges(pF), 1.0)
    np.savez_compressed(os.path.join(HERE, "lm_sphere_n24.npz"), src=pV, src_f=pF, tgt=qV, tgt_f=qF, N=M,
                        x0=x0, x=x, final_cost=st["final_cost"], iterations=st["iterations"])
    # a pair above the register/LDS solver's 8192 vertices (icosphere level 5
    # source, level 4 target, 48^3): only the converged energy is kept, since
    # the oracle's exact-step LM needs about a minute for it
    pV, pF, qV, qF = synth.displaced_sphere_pair(31, src_level=5, tgt_level=4)
    qn, qs, qp = O.normalize(qV)
    M = 48
    g, _, _ = O.build_grid_bvh(qn, qF, M)
    x0 = O.apply_transform(pV, qs, qp)
    _, st, _ = O.lm_solve(g, M, x0, O.face_edges(pF), 1.0)
    np.savez_compressed(os.path.join(HERE, "lm_sphere_l5_n48.npz"), seed=31, src_level=5, tgt_level=4, N=M, lam=1.0,
                        final_cost=st["final_cost"], iterations=st["iterations"])
    for f in sorted(os.listdir(HERE)):
        if f.endswith(".npz"):
            print(f, os.path.getsize(os.path.join(HERE, f)))


def cfg4_l6():
    """Config 4's largest pair: icosphere level 6 source (40962 vertices)
    against a displaced level-6 target at 128^3, lambda 1.3, the oracle's
    converged energy (its exact-step LM needs tens of minutes here)."""
    seed, level, M, lam = 6000, 6, 128, 1.3
    pV, pF, qV, qF = synth.icosphere_pair(seed, level)
    qn, qs, qp = O.normalize(qV)
    g, _, _ = O.build_grid_bvh(qn, qF, M)
    x0 = O.apply_transform(pV, qs, qp)
    _, st, _ = O.lm_solve(g, M, x0, O.face_edges(pF), lam)
    np.savez_compressed(os.path.join(HERE, "lm_cfg4_l6_n128.npz"), seed=seed, level=level, N=M, lam=lam,
                        final_cost=st["final_cost"], initial_cost=st["initial_cost"], iterations=st["iterations"])
    print("lm_cfg4_l6_n128", st)


FIXTURES = {"all": main, "cfg4_l6": cfg4_l6}

if __name__ == "__main__":
    for name in sys.argv[1:] or ["all"]:
        FIXTURES[name]()
