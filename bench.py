#!/usr/bin/env python3
"""Deformation throughput on one MI355X.

Workload (one "step"): the rigid_deform pipeline (rigid_deform.cc:49-61) over
a batch of B source/target pairs resident in HBM — per pair: normalize the
target (Mesh::Normalize), build its 64^3 distance grid
(Mesh::ConstructDistanceField), bring the source into the target frame, solve
Deformer::Deform (Levenberg-Marquardt, <= 100 iterations, lambda = 1) and write
the result in the target frame.  Pairs are the seeded 12-box unions of
5000 vertices against their jittered targets (meshode_amd/synth.py).

Reports pairs/s (whole batch), p50/p99 per-pair latency inside the batch, the
one-pair-at-a-time GPU rate, the LM kernel's roofline position, the CPU
oracle (FP64 restatement of the reference pipeline, 1 core) on the same
pair, and the energy agreement of that pair with the oracle.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0


def lm_algorithmic_bytes(nv, nh, lm_iters, pcg_iters):
    """Compulsory traffic of the LM solve (FP32 vectors, int32 indices):
    per LM iteration 96 V + 44 H bytes (lookup + gradient, edge residuals and
    J^T r, update), per PCG iteration 184 V + 4 H bytes (matrix-free product,
    preconditioner apply, fused axpy/dot passes)."""
    return lm_iters * (96 * nv + 44 * nh) + pcg_iters * (184 * nv + 4 * nh)


def grid_algorithmic_bytes(N, nf_t):
    return 4 * N ** 3 + 36 * nf_t


def cpu_baseline(pair, N, lam):
    """The oracle (FP64 restatement of the reference pipeline: Normalize,
    AABB-tree grid build, ApplyTransform, Ceres-restated LM with exact
    sparse Cholesky solves) on one pair, single thread."""
    from oracle import oracle as O
    sV, sF, tV, tF = pair
    t0 = time.perf_counter()
    tn, sc, pos = O.normalize(tV)
    d, _, _ = O.build_grid_bvh(tn, tF, N)
    sn = O.apply_transform(sV, sc, pos)
    x, st, _ = O.lm_solve(d, N, sn, O.face_edges(sF), lam)
    dt = time.perf_counter() - t0
    return dt, st, d, sn, sc, pos


def main():
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--pairs", type=int, default=512, help="pairs per step (batch)")
    ap.add_argument("--distinct", type=int, default=128, help="distinct seeded pairs (repeated to fill the batch)")
    ap.add_argument("--grid", type=int, default=64)
    ap.add_argument("--verts", type=int, default=5000)
    ap.add_argument("--lam", type=float, default=1.0)
    ap.add_argument("--plain-pairs", type=int, default=8, help="pairs run one at a time for the plain GPU rate")
    ap.add_argument("--block-threads", type=int, default=1024, help="threads per pair of the global-memory kernel")
    ap.add_argument("--kernel-variant", type=int, default=0, help="0: register/LDS kernel when a pair fits; 1: global")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    args = ap.parse_args()

    if args.gpus != 1:
        print(f"bench.py: this build targets one GPU; --gpus {args.gpus} is not supported "
              "(pairs are independent: run one process per GPU).", file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import torch

    from meshode_amd import _lib, deform, synth

    assert torch.cuda.is_available(), "bench.py needs a GPU"
    torch.cuda.set_device(0)
    _lib.load()

    t0 = time.perf_counter()
    distinct = [synth.box_union_pair(s, n_verts=args.verts) for s in range(min(args.distinct, args.pairs))]
    pairs = [distinct[i % len(distinct)] for i in range(args.pairs)]
    gen_s = time.perf_counter() - t0
    print(f"[bench] generated {len(distinct)} distinct pairs in {gen_s:.1f}s; batch {args.pairs}", file=sys.stderr,
          flush=True)

    batch = deform.PairBatch(pairs, grid_resolution=args.grid, lambdas=args.lam)
    opt = deform.default_options(block_threads=args.block_threads, timing=1, kernel_variant=args.kernel_variant)
    torch.cuda.synchronize()

    for w in range(args.warmup):
        batch.run(opt)
        print(f"[bench] warmup {w + 1}/{args.warmup} done", file=sys.stderr, flush=True)

    torch.cuda.synchronize()
    stats_all, timings = [], []
    t_start = time.perf_counter()
    for k in range(args.steps):
        st = batch.run(opt)
        stats_all.append(st)
        timings.append(_lib.last_timings())
        print(f"[bench] step {k + 1}/{args.steps}: {timings[-1]['total_ms']:.1f} ms", file=sys.stderr, flush=True)
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t_start

    value = args.pairs * args.steps / elapsed
    ms_per_step = 1e3 * elapsed / args.steps

    # per-pair latency inside the batch: completion time since the call began
    lat = []
    for st, tm in zip(stats_all, timings):
        for s in st:
            lat.append((s["clock_end"] - tm["clock_start"]) / tm["clock_hz"] * 1e3)
    lat = np.array(lat)

    # roofline of the dominant kernel (LM solver), algorithmic bytes per launch
    nv = np.array([len(p[0]) for p in pairs])
    nh = np.array([3 * len(p[1]) for p in pairs])
    lm_bytes = []
    for st in stats_all:
        it = np.array([s["iterations"] for s in st])
        pc = np.array([s["pcg_iterations"] for s in st])
        lm_bytes.append(float(np.sum(lm_algorithmic_bytes(nv, nh, it, pc))))
    lm_ms = np.array([t["lm_ms"] for t in timings])
    achieved = float(np.mean(lm_bytes) / (lm_ms.mean() / 1e3) / 1e9)
    last = stats_all[-1]

    # plain path: one pair per call, sequentially
    plain = deform.PairBatch(distinct[: args.plain_pairs], grid_resolution=args.grid, lambdas=args.lam)
    singles = [deform.PairBatch([p], grid_resolution=args.grid, lambdas=args.lam) for p in distinct[: args.plain_pairs]]
    singles[0].run(opt)
    torch.cuda.synchronize()
    tp = time.perf_counter()
    for b in singles:
        b.run(opt)
    torch.cuda.synchronize()
    plain_rate = len(singles) / (time.perf_counter() - tp)
    del plain

    result = {
        "metric": "deformation pairs/sec (5k verts, 64^3 grid, energy tol 1e-3)",
        "value": value,
        "unit": "pairs/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {
            "workload": "rigid_deform pipeline, batch of seeded 12-box unions (5000 verts) vs jittered targets",
            "pairs_per_step": args.pairs,
            "distinct_pairs": len(distinct),
            "source_verts": args.verts,
            "grid": args.grid,
            "lambda": args.lam,
            "lm_max_iterations": 100,
            "block_threads": args.block_threads,
            "kernel_variant": args.kernel_variant,
            "lm_shapes": os.environ.get("MO_LM_SHAPES", "default"),
        },
        "latency_ms": {"p50": float(np.percentile(lat, 50)), "p99": float(np.percentile(lat, 99)),
                       "max": float(lat.max())},
        "plain_one_pair_per_call_pairs_per_s": plain_rate,
        "batched_over_plain": value / plain_rate,
        "phase_ms_mean": {k: float(np.mean([t[k] for t in timings])) for k in ("grid_ms", "setup_ms", "lm_ms",
                                                                                 "total_ms")},
        "lm_iterations_mean": float(np.mean([s["iterations"] for s in last])),
        "pcg_iterations_per_lm_iteration": float(np.sum([s["pcg_iterations"] for s in last]) /
                                                 max(1, np.sum([s["iterations"] for s in last]))),
        "roofline": {
            "kernel": "k_lm_solve",
            "bound": "hbm",
            "achieved": achieved,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS,
            "traffic": None,
            "algorithmic_bytes_per_launch": float(np.mean(lm_bytes)),
        },
    }

    if not args.no_cpu_baseline:
        dt, so, grid, sn, sc, pos = cpu_baseline(distinct[0], args.grid, args.lam)
        from oracle import oracle as O
        xg = (batch.out[0].cpu().numpy() - pos) / sc
        eg, _, _ = O.lm_cost(grid, args.grid, xg, sn, O.face_edges(distinct[0][1]), args.lam)
        result["cpu_baseline"] = {
            "value": 1.0 / dt,
            "unit": "pairs/s",
            "cores": 1,
            "kind": "port",
            "sample": "pair 0 of the workload (5000 verts, 64^3): normalize + AABB-tree grid + "
                      "Ceres-restated LM with exact sparse Cholesky, FP64, one thread",
            "seconds": dt,
            "host_cpu": _cpu_model(),
            "host_nproc": os.cpu_count(),
        }
        result["energy_check_pair0"] = {"oracle": so["final_cost"], "gpu": eg,
                                        "rel_diff": (eg - so["final_cost"]) / so["final_cost"],
                                        "oracle_iterations": so["iterations"],
                                        "gpu_iterations": stats_all[-1][0]["iterations"]}
    print(json.dumps(result), flush=True)


def _cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


if __name__ == "__main__":
    main()
