#!/usr/bin/env python3
"""Deformation throughput on one MI355X.

Workload (one "step"): the rigid_deform pipeline (rigid_deform.cc:49-61) over
a batch of B source/target pairs resident in HBM — per pair: normalize the
target (Mesh::Normalize), build its 64^3 distance grid
(Mesh::ConstructDistanceField), bring the source into the target frame, solve
Deformer::Deform (Levenberg-Marquardt, <= 100 iterations, lambda = 1) and write
the result in the target frame.  Pairs are the seeded 12-box unions of
5000 vertices against their jittered targets (meshode_amd/synth.py).

Reports pairs/s over K steps with two batches in flight on two HIP streams
(mo_rigid_deform_batch_async: one batch's grid builds fill the CUs that the
previous batch's last solves leave idle), the same steps one at a time on one
stream (with per-phase times), p50/p99 per-pair latency inside the batch, the
one-pair-at-a-time GPU rate, the LM kernel's roofline position, the CPU
oracle (FP64 restatement of the reference pipeline, 1 core) on the same
pair, and the energy agreement of that pair with the oracle.
"""
import argparse
import collections
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
FP64_VALU_PEAK_TFLOPS = 78.6    # MI355X FP64 vector (FMA = 2 flops)
FP32_VALU_PEAK_TFLOPS = 157.3   # MI355X FP32 vector (published; packed FP32)

# Flop model of the register/LDS LM kernel (FP32 vector work per vertex):
# per PCG iteration 93 + 6 E_v (matrix-free A z with E_v adjacency entries
# per vertex at 3 FMAs each, recurrences, the residual recovered as M z,
# Sherman-Morrison preconditioner, three dot products); per LM iteration
# 150 + 20 E_v (distance lookup and gradient, edge energy in the Jacobian
# and candidate evaluations).  rocprofv3's SQ_INSTS_VALU_FLOPS_FP32/FP64 on
# the headline launch read 6.4e11 flops against this model's 6.4e11.
LM_PCG_FLOPS_BASE, LM_PCG_FLOPS_PER_ENTRY = 93, 6
LM_ITER_FLOPS_BASE, LM_ITER_FLOPS_PER_ENTRY = 150, 20


def lm_algorithmic_flops(nv, entries, lm_iters, pcg_iters):
    """Flops of one pair's LM solve: nv vertices, `entries` adjacency
    entries (distinct neighbours summed over vertices)."""
    import numpy as np
    ev = np.asarray(entries, dtype=np.float64) / np.maximum(nv, 1)
    return nv * (pcg_iters * (LM_PCG_FLOPS_BASE + LM_PCG_FLOPS_PER_ENTRY * ev) +
                 lm_iters * (LM_ITER_FLOPS_BASE + LM_ITER_FLOPS_PER_ENTRY * ev))


# Flop model of the grid build (FP64, no FMA contraction): 60 flops per
# point-triangle distance evaluation (face-region test, or the three
# canonical edge distances) and 25 per face bound check (point-to-AABB and
# plane bounds); both counts come live from the kernel (mo_last_grid_stats).
# rocprofv3's SQ_INSTS_VALU_FLOPS_FP64 on the headline launch reads 9.3e11
# flops against this model's ~7e11 (the rest: tile scans and seeding).
GRID_FLOPS_PER_EXACT, GRID_FLOPS_PER_BOUND = 60, 25


def adjacency_entries(F):
    """Distinct undirected neighbours summed over vertices (the LM kernel's
    adjacency entries) of a triangle mesh."""
    import numpy as np
    a = np.concatenate([F[:, 0], F[:, 1], F[:, 2], F[:, 1], F[:, 2], F[:, 0]]).astype(np.int64)
    b = np.concatenate([F[:, 1], F[:, 2], F[:, 0], F[:, 0], F[:, 1], F[:, 2]]).astype(np.int64)
    keep = a != b
    return int(len(np.unique(a[keep] * (int(F.max()) + 1) + b[keep])))


# HBM bytes per launch measured with rocprofv3 counter passes (FETCH_SIZE x
# 1.99975 and WRITE_SIZE x 1.0, the factors a plain 256 MiB copy reads) on
# the default workload; keyed by (kernel, pairs, distinct pairs, source
# verts, grid); other workloads report null.
MEASURED_TRAFFIC = {("k_lm_solve_lds", 512, 128, 5000, 64): 28.03e9,
                    ("k_grid_build", 512, 128, 5000, 64): 14.56e9}
# Wavefront VALU instructions per launch (rocprofv3 SQ_INSTS_VALU, same
# workload and key): with the live launch time they give the fraction of the
# SIMDs' issue cycles the kernel's VALU work fills (a wave64 VALU
# instruction occupies a SIMD 4 cycles; 1024 SIMDs at 2.4 GHz), which is
# what bounds both kernels.
MEASURED_VALU_INSTS = {("k_lm_solve_lds", 512, 128, 5000, 64): 1.252e10,
                       ("k_grid_build", 512, 128, 5000, 64): 1.565e10}
SIMDS, ENGINE_HZ = 1024, 2.4e9


def valu_issue_frac(kernel, wkey, launch_ms):
    n = MEASURED_VALU_INSTS.get((kernel,) + wkey)
    return None if n is None else n * 4 / (SIMDS * launch_ms * 1e-3 * ENGINE_HZ)


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


def cpu_baseline_all_cores(pairs, N, lam):
    """The same oracle pipeline on one pair per host core at once (threads:
    the oracle's C entry points release the GIL), aggregate pairs/s."""
    from concurrent.futures import ThreadPoolExecutor
    cores = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else os.cpu_count()
    work = [pairs[i % len(pairs)] for i in range(cores)]
    t0 = time.perf_counter()
    with ThreadPoolExecutor(max_workers=cores) as ex:
        list(ex.map(lambda p: cpu_baseline(p, N, lam), work))
    dt = time.perf_counter() - t0
    return {"value": len(work) / dt, "unit": "pairs/s", "cores": cores, "kind": "port",
            "sample": "%d pairs of the workload, one per core in parallel (same oracle pipeline)" % len(work),
            "seconds": dt}


def main():
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--pairs", type=int, default=512, help="pairs per step (batch)")
    ap.add_argument("--distinct", type=int, default=128, help="distinct seeded pairs (repeated to fill the batch)")
    ap.add_argument("--grid", type=int, default=64)
    ap.add_argument("--verts", type=int, default=5000)
    ap.add_argument("--lam", type=float, default=1.0)
    ap.add_argument("--plain-pairs", type=int, default=8, help="pairs run one at a time for the plain GPU rate")
    ap.add_argument("--block-threads", type=int, default=1024, help="threads per pair of the global-memory kernel")
    ap.add_argument("--kernel-variant", type=int, default=0,
                    help="0: register/LDS kernel when a pair fits, else teams; 1: one-workgroup global; 2: teams")
    ap.add_argument("--pcg-tol", type=float, default=None, help="inner PCG relative tolerance (library default if unset)")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-graph", action="store_true", help="skip the captured-graph replay measurement")
    ap.add_argument("--no-autotune", action="store_true",
                    help="skip the launch-shape autotuner (untimed, before warm-up) and use the built-in order")
    ap.add_argument("--mode", default="batch", choices=["batch", "serve", "serve5k", "adam", "ode", "flow256", "neural2", "rot", "cfg1", "cfg4"],
                    help="batch: the headline line; serve / adam / ode: the serving, pyDeform-Adam and "
                         "NeuralODE workloads (their own JSON line)")
    ap.add_argument("--duration", type=float, default=None, help="serve: seconds of Poisson arrivals (default 60)")
    ap.add_argument("--serve-pairs", type=int, default=96, help="serve: distinct (source, target) pairs in the pool")
    ap.add_argument("--budget-gb", type=float, default=200.0, help="serve: template cache / admission budget")
    ap.add_argument("--rate", type=float, default=200.0, help="serve: requests per second")
    ap.add_argument("--max-verts", type=int, default=100000, help="serve: largest source size")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="first run one plain 256 MiB device copy (known bytes) to scale HBM counters")
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
    if args.copy_calibration:
        a = torch.ones(64 * 1024 * 1024, dtype=torch.float32, device="cuda")
        b = torch.empty_like(a)
        torch.cuda.synchronize()
        b.copy_(a)
        torch.cuda.synchronize()
        del a, b
    if args.mode != "batch":
        print(json.dumps({"serve": bench_serve, "serve5k": bench_serve5k, "adam": bench_adam, "ode": bench_ode, "flow256": bench_flow256,
                          "neural2": bench_neural2, "rot": bench_rot, "cfg1": bench_cfg1,
                          "cfg4": bench_cfg4}[args.mode](args)),
              flush=True)
        return

    t0 = time.perf_counter()
    distinct = [synth.box_union_pair(s, n_verts=args.verts) for s in range(min(args.distinct, args.pairs))]
    pairs = [distinct[i % len(distinct)] for i in range(args.pairs)]
    gen_s = time.perf_counter() - t0
    print(f"[bench] generated {len(distinct)} distinct pairs in {gen_s:.1f}s; batch {args.pairs}", file=sys.stderr,
          flush=True)

    batch = deform.PairBatch(pairs, grid_resolution=args.grid, lambdas=args.lam)
    opt = deform.default_options(block_threads=args.block_threads, timing=1, kernel_variant=args.kernel_variant)
    if args.pcg_tol is not None:
        opt.pcg_relative_tolerance = args.pcg_tol
    shapes_used = _autotune(args, distinct, args.grid, args.lam, opt)
    torch.cuda.synchronize()

    for w in range(args.warmup):
        batch.run(opt)
        print(f"[bench] warmup {w + 1}/{args.warmup} done", file=sys.stderr, flush=True)

    # the headline: K steps, each one batch of args.pairs pairs, with two
    # batches in flight on two streams (mo_rigid_deform_batch_async); step
    # k+1's call returns once its solver launches are queued, and step k is
    # waited for after that, so its last solves overlap the next grid builds
    aopt = deform.default_options(block_threads=args.block_threads, timing=0, kernel_variant=args.kernel_variant)
    if args.pcg_tol is not None:
        aopt.pcg_relative_tolerance = args.pcg_tol
    pipe = [deform.PairBatch(pairs, grid_resolution=args.grid, lambdas=args.lam) for _ in range(2)]
    streams = [torch.cuda.Stream(), torch.cuda.Stream()]
    for j in range(2):
        with torch.cuda.stream(streams[j]):
            pipe[j].run_async(aopt)
        pipe[j].wait()
    torch.cuda.synchronize()
    pipe_stats = []
    t_start = time.perf_counter()
    prev = None
    for k in range(args.steps):
        with torch.cuda.stream(streams[k % 2]):
            pipe[k % 2].run_async(aopt, after=pipe[prev] if prev is not None else None)
        if prev is not None:
            pipe_stats.append(pipe[prev].wait())
        prev = k % 2
    pipe_stats.append(pipe[prev].wait())
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t_start
    value = args.pairs * args.steps / elapsed
    ms_per_step = 1e3 * elapsed / args.steps
    print(f"[bench] pipelined: {ms_per_step:.1f} ms/step, {value:.0f} pairs/s", file=sys.stderr, flush=True)

    # the same steps one at a time on one stream, with HIP events around the
    # phases (grid, set-up, LM)
    torch.cuda.synchronize()
    stats_all, timings = [], []
    t1 = time.perf_counter()
    for k in range(args.steps):
        st = batch.run(opt)
        stats_all.append(st)
        timings.append(_lib.last_timings())
        print(f"[bench] step {k + 1}/{args.steps}: {timings[-1]['total_ms']:.1f} ms", file=sys.stderr, flush=True)
    torch.cuda.synchronize()
    el1 = time.perf_counter() - t1
    pipe_equal = all(torch.equal(a, b) for pb in pipe for a, b in zip(pb.out, batch.out))
    one_stream = {"pairs_per_s": args.pairs * args.steps / el1, "ms_per_step": 1e3 * el1 / args.steps,
                  "outputs_equal_pipelined": bool(pipe_equal)}

    # the same step as a captured HIP graph (mo_rigid_deform_batch_graph):
    # every kernel of the chain replays on the same inputs; timed like the
    # eager steps (timing events are not captured, so opt.timing is 0)
    graph_line = None
    if not args.no_graph:
        gopt = deform.default_options(block_threads=args.block_threads, timing=0,
                                      kernel_variant=args.kernel_variant)
        if args.pcg_tol is not None:
            gopt.pcg_relative_tolerance = args.pcg_tol
        gbatch = deform.PairBatch(pairs, grid_resolution=args.grid, lambdas=args.lam)
        gst = gbatch.capture(gopt)
        gbatch.replay()
        torch.cuda.synchronize()
        tg = time.perf_counter()
        for k in range(args.steps):
            gst = gbatch.replay()
        torch.cuda.synchronize()
        gel = time.perf_counter() - tg
        same = all(torch.equal(a, b) for a, b in zip(gbatch.out, batch.out))
        graph_line = {"pairs_per_s": args.pairs * args.steps / gel, "ms_per_step": 1e3 * gel / args.steps,
                      "outputs_equal_eager": bool(same),
                      "lm_iterations_mean": float(np.mean([s["iterations"] for s in gst]))}
        gbatch.release()
        print(f"[bench] graph replay: {graph_line['ms_per_step']:.1f} ms/step, outputs equal eager: {same}",
              file=sys.stderr, flush=True)
    # one more step, after the timed region, with the grid build's work
    # counters on (timing 2: their atomics cost a few percent of the phase)
    opt.timing = 2
    batch.run(opt)
    opt.timing = 1

    # per-pair latency inside the batch: completion time since the call began
    lat = []
    for st, tm in zip(stats_all, timings):
        for s in st:
            lat.append((s["clock_end"] - tm["clock_start"]) / tm["clock_hz"] * 1e3)
    lat = np.array(lat)

    # rooflines: the grid build (FP64 VALU, work counted live by the
    # kernel) and the LM solver (FP32 VALU, algorithmic flop model); the one
    # with the longer live time is the dominant kernel
    nv = np.array([len(p[0]) for p in pairs])
    ent_d = [adjacency_entries(p[1]) for p in distinct]
    ent = np.array([ent_d[i % len(distinct)] for i in range(len(pairs))])
    lm_flops = []
    for st in stats_all:
        it = np.array([s["iterations"] for s in st])
        pc = np.array([s["pcg_iterations"] for s in st])
        lm_flops.append(float(np.sum(lm_algorithmic_flops(nv, ent, it, pc))))
    lm_ms = float(np.mean([t["lm_ms"] for t in timings]))
    grid_ms = float(np.mean([t["grid_ms"] for t in timings]))
    gs = _lib.last_grid_stats()
    grid_flops = GRID_FLOPS_PER_EXACT * gs["exact_tests"] + GRID_FLOPS_PER_BOUND * gs["bound_checks"]
    lm_tf = float(np.mean(lm_flops)) / (lm_ms / 1e3) / 1e12
    grid_tf = grid_flops / (grid_ms / 1e3) / 1e12
    cells = int(sum(args.grid ** 3 for _ in pairs))
    wkey = (args.pairs, len(distinct), args.verts, args.grid)
    roof_grid = {"kernel": "k_grid_build", "bound": "fp64-valu", "achieved": grid_tf, "peak": FP64_VALU_PEAK_TFLOPS,
                 "unit": "TFLOP/s", "frac": grid_tf / FP64_VALU_PEAK_TFLOPS,
                 "traffic": MEASURED_TRAFFIC.get(("k_grid_build",) + wkey),
                 "valu_issue_frac": valu_issue_frac("k_grid_build", wkey, grid_ms),
                 "flops_per_launch": grid_flops, "launch_ms": grid_ms,
                 "exact_tests_per_cell": gs["exact_tests"] / cells,
                 "exact_tests_needed_per_cell": gs["exact_tests_needed"] / cells,
                 "bound_checks_per_cell": gs["bound_checks"] / cells,
                 "tiles_per_64_cells": 64 * gs["tiles"] / max(1, cells),
                 "bound_check_flop_share": GRID_FLOPS_PER_BOUND * gs["bound_checks"] / max(grid_flops, 1)}
    roof_lm = {"kernel": "k_lm_solve_lds", "bound": "fp32-valu", "achieved": lm_tf, "peak": FP32_VALU_PEAK_TFLOPS,
               "unit": "TFLOP/s", "frac": lm_tf / FP32_VALU_PEAK_TFLOPS,
               "traffic": MEASURED_TRAFFIC.get(("k_lm_solve_lds",) + wkey),
               "valu_issue_frac": valu_issue_frac("k_lm_solve_lds", wkey, lm_ms),
               "flops_per_launch": float(np.mean(lm_flops)), "launch_ms": lm_ms}
    assert roof_grid["frac"] <= 1.0 and roof_lm["frac"] <= 1.0, (roof_grid, roof_lm)
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
            "batches_in_flight": 2,
            "pairs_per_step": args.pairs,
            "distinct_pairs": len(distinct),
            "source_verts": args.verts,
            "grid": args.grid,
            "lambda": args.lam,
            "lm_max_iterations": 100,
            "block_threads": opt.block_threads,
            "kernel_variant": args.kernel_variant,
            "lm_shapes": shapes_used,
        },
        "latency_ms": {"p50": float(np.percentile(lat, 50)), "p99": float(np.percentile(lat, 99)),
                       "max": float(lat.max())},
        "plain_one_pair_per_call_pairs_per_s": plain_rate,
        "batched_over_plain": value / plain_rate,
        "one_stream": one_stream,
        "phase_ms_mean": {k: float(np.mean([t[k] for t in timings])) for k in ("grid_ms", "setup_ms", "lm_ms",
                                                                                 "total_ms")},
        "lm_iterations_mean": float(np.mean([s["iterations"] for s in last])),
        "pcg_iterations_per_lm_iteration": float(np.sum([s["pcg_iterations"] for s in last]) /
                                                 max(1, np.sum([s["iterations"] for s in last]))),
        "roofline": roof_grid if grid_ms >= lm_ms else roof_lm,
        "roofline_grid": roof_grid,
        "roofline_lm": roof_lm,
        "graph_replay": graph_line,
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
        result["cpu_baseline_all_cores"] = cpu_baseline_all_cores(distinct, args.grid, args.lam)
        result["energy_check_pair0"] = {"oracle": so["final_cost"], "gpu": eg,
                                        "rel_diff": (eg - so["final_cost"]) / so["final_cost"],
                                        "oracle_iterations": so["iterations"],
                                        "gpu_iterations": stats_all[-1][0]["iterations"]}
    print(json.dumps(result), flush=True)


def bench_cfg1(args):
    """Config 1: icosphere L4 sources (2562 vertices) against displaced L5
    icospheres scaled (1, 0.7, 1.3) (10242 vertices), 64^3, lambda 1; the
    rigid_deform pipeline over a batch of args.pairs pairs (args.distinct
    distinct seeds), pairs/s; the oracle on pair 0 for the energy check and
    the CPU baseline."""
    import numpy as np
    import torch
    from meshode_amd import _lib, deform, synth
    distinct = [synth.icosphere_pair(s, 4) for s in range(min(args.distinct, args.pairs))]
    pairs = [distinct[i % len(distinct)] for i in range(args.pairs)]
    batch = deform.PairBatch(pairs, grid_resolution=args.grid, lambdas=args.lam)
    opt = deform.default_options(timing=1)
    shapes_used = _autotune(args, distinct, args.grid, args.lam, opt)
    for _ in range(max(args.warmup, 0)):
        batch.run(opt)
    torch.cuda.synchronize()
    steps = max(args.steps, 1)
    t0 = time.perf_counter()
    for _ in range(steps):
        st = batch.run(opt)
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    tm = _lib.last_timings()
    out = {"metric": "deformation pairs/sec (config 1: 2562 -> 10242 verts, 64^3)", "value": args.pairs * steps / elapsed,
           "unit": "pairs/s", "n_gpus": 1, "steps": steps, "warmup": args.warmup,
           "ms_per_step": 1e3 * elapsed / steps, "higher_is_better": True, "dtype": "f32 (LM), f64 (grid)",
           "data": "synthetic", "phase_ms_last": {k: tm[k] for k in ("grid_ms", "setup_ms", "lm_ms", "total_ms")},
           "config": {"workload": "config 1: icosphere L4 vs displaced icosphere L5 (x 1, 0.7, 1.3)",
                      "pairs": args.pairs, "distinct": len(distinct), "grid": args.grid, "lambda": args.lam,
                      "lm_shapes": shapes_used}}
    if not args.no_cpu_baseline:
        dt, so, grid, sn, sc, pos = cpu_baseline(distinct[0], args.grid, args.lam)
        from oracle import oracle as O
        eg, _, _ = O.lm_cost(grid, args.grid, (batch.out[0].cpu().numpy() - pos) / sc, sn,
                             O.face_edges(distinct[0][1]), args.lam)
        out["cpu_baseline"] = {"value": 1.0 / dt, "unit": "pairs/s", "cores": 1, "kind": "port",
                               "sample": "pair 0 (oracle pipeline, FP64, one thread)", "seconds": dt}
        out["energy_check_pair0"] = {"oracle": so["final_cost"], "gpu": eg,
                                     "rel_diff": (eg - so["final_cost"]) / so["final_cost"]}
    return out


def _autotune(args, pairs, grid, lam, opt, repeats=2):
    """Run the launch-shape autotuner on the workload's pairs before the
    warm-up (untimed) and install its table; returns the shape summary for
    the JSON line.  Sets opt.block_threads when the workload has pairs for
    the global-memory kernels (and opt.team_vertices)."""
    if args.no_autotune or args.kernel_variant != 0 or "MO_LM_SHAPES" in os.environ:
        return os.environ.get("MO_LM_SHAPES", "built-in order")
    from meshode_amd import autotune
    t0 = time.perf_counter()
    res = autotune.tune(pairs, grid_resolution=grid, lambdas=lam, options=opt, repeats=repeats,
                        max_pairs=len(pairs))
    if res["block_threads"]:
        opt.block_threads = res["block_threads"]
        opt.team_vertices = res["team_vertices"]
    desc = autotune.describe(res)
    print(f"[bench] autotuned launch shapes in {time.perf_counter() - t0:.1f}s: {desc}", file=sys.stderr, flush=True)
    return "autotuned: " + desc


def bench_cfg4(args):
    """Config 4: 256 pairs in one rigid_deform batch, each built like config 1
    (displaced-icosphere target one level finer, capped at 6) with the source
    level drawn from {3, 4, 5, 6} (642 to 40962 vertices), the grid from
    {64^3, 128^3} and lambda ~ U(0.5, 2) (numpy PCG64 seed 4); fixed work of
    50 LM iterations with 30 PCG iterations each (tolerances zeroed).  Pairs
    of at most 8192 vertices run on the register/LDS kernel, larger ones on
    the global-memory kernel, in the same call.  8 distinct seeds per level
    are repeated to fill the batch (the level-6 generator takes ~1 s a pair
    on the host).  Launch shapes come from the autotuner (untimed)."""
    import numpy as np
    import torch
    from meshode_amd import _lib, deform, synth
    n = 256
    rng = np.random.Generator(np.random.PCG64(4))
    levels = rng.choice([3, 4, 5, 6], n)
    grids = [int(g) for g in rng.choice([64, 128], n)]
    lams = [float(x) for x in rng.uniform(0.5, 2.0, n)]
    cache = {}
    pairs = []
    for i in range(n):
        key = (int(levels[i]), i % 8)
        if key not in cache:
            cache[key] = synth.icosphere_pair(1000 * key[0] + key[1], key[0])
        pairs.append(cache[key])
    batch = deform.PairBatch(pairs, grid_resolution=grids, lambdas=lams)
    opt = deform.default_options(timing=1, max_num_iterations=50, pcg_max_iterations=30,
                                 function_tolerance=0.0, gradient_tolerance=0.0, parameter_tolerance=0.0,
                                 pcg_relative_tolerance=0.0, block_threads=args.block_threads)
    shapes_used = _autotune(args, pairs, grids, lams, opt, repeats=1)
    for _ in range(max(args.warmup, 0)):
        batch.run(opt)
    torch.cuda.synchronize()
    steps = max(args.steps, 1)
    t0 = time.perf_counter()
    for _ in range(steps):
        st = batch.run(opt)
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    tm = _lib.last_timings()
    nv = np.array([len(p[0]) for p in pairs])
    it = np.array([s["iterations"] for s in st])
    pcg = np.array([s["pcg_iterations"] for s in st])
    return {"metric": "deformation pairs/sec (config 4: 256 mixed pairs, L3-L6, 64^3/128^3)",
            "value": n * steps / elapsed, "unit": "pairs/s", "n_gpus": 1, "steps": steps, "warmup": args.warmup,
            "ms_per_step": 1e3 * elapsed / steps, "higher_is_better": True, "dtype": "f32 (LM), f64 (grid)",
            "data": "synthetic", "phase_ms_last": {k: tm[k] for k in ("grid_ms", "setup_ms", "lm_ms", "total_ms")},
            "lm_iterations": {"min": int(it.min()), "max": int(it.max())},
            "pcg_per_lm_iteration": float(pcg.sum() / max(it.sum(), 1)),
            "vertices_per_step": int(nv.sum()),
            "config": {"workload": "config 4: icosphere L3-L6 sources vs displaced icospheres one level finer",
                       "pairs": n, "levels": {str(L): int((levels == L).sum()) for L in (3, 4, 5, 6)},
                       "grids": {str(g): grids.count(g) for g in (64, 128)}, "lambda": "U(0.5, 2)",
                       "lm_iterations": 50, "pcg_iterations": 30, "seed": 4, "lm_shapes": shapes_used,
                       "block_threads": opt.block_threads, "team_vertices": opt.team_vertices}}


def _serve_pair(job):
    """One (source, target) pair of the serving pool: config-1-like
    (icosphere source, displaced icosphere target one level finer, capped at
    6) or config-2-like (box union of exactly nv vertices vs its jittered
    boxes); run in a worker process."""
    kind, seed, nv = job
    from meshode_amd import synth
    if kind == 0:
        lvl = 4 if nv < 6000 else 5 if nv < 25000 else 6
        return synth.icosphere_pair(seed, lvl)
    return synth.box_union_pair(seed, n_verts=max(1000, nv - nv % 2))


def bench_serve(args):
    """Config 5: Poisson arrivals at args.rate requests/s for args.duration
    seconds through the continuous-batching server, cache starting cold.
    A pool of args.serve_pairs distinct pairs (half built like config 1, half
    like config 2; source sizes log-uniform 2k-args.max_verts vertices; each
    pair's grid drawn from {64, 128, 256}^3) is generated on the host first;
    every request takes a pool pair uniformly at random and moves its source
    by a fresh random rotation (up to 3 degrees), scale (+-2 %) and shift
    (+-1 % of the extent), so no two requests are the same problem.  The
    first request for a target builds its template (normalization + grid);
    the LRU cache works under args.budget_gb.  Reports pairs/s, p50/p99
    latency (all requests and cache misses), builds and evictions."""
    import numpy as np
    from concurrent.futures import ProcessPoolExecutor
    from meshode_amd import serve, synth
    rng = np.random.default_rng(7)
    n_pool = args.serve_pairs
    sizes = np.exp(rng.uniform(np.log(2000), np.log(args.max_verts), n_pool)).astype(int)
    kinds = np.arange(n_pool) % 2
    grids = [int(g) for g in rng.choice([64, 128, 256], n_pool)]
    t0 = time.perf_counter()
    workers = min(16, len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else os.cpu_count())
    with ProcessPoolExecutor(max_workers=workers) as ex:
        pool = list(ex.map(_serve_pair, [(int(k), 5000 + i, int(n)) for i, (k, n) in enumerate(zip(kinds, sizes))]))
    print(f"[bench] serve: {n_pool} pool pairs generated in {time.perf_counter() - t0:.1f}s", file=sys.stderr,
          flush=True)
    duration = args.duration if args.duration is not None else 60.0
    n_req = int(args.rate * duration)
    gaps = rng.exponential(1.0 / args.rate, n_req)
    picks = rng.integers(0, n_pool, n_req)
    ang = np.radians(rng.uniform(-3, 3, (n_req, 3)))
    scl = rng.uniform(0.98, 1.02, n_req)
    shf = rng.uniform(-0.01, 0.01, (n_req, 3))

    def moved(V, k):
        a, b, c = ang[k]
        Rx = np.array([[1, 0, 0], [0, np.cos(a), -np.sin(a)], [0, np.sin(a), np.cos(a)]])
        Ry = np.array([[np.cos(b), 0, np.sin(b)], [0, 1, 0], [-np.sin(b), 0, np.cos(b)]])
        Rz = np.array([[np.cos(c), -np.sin(c), 0], [np.sin(c), np.cos(c), 0], [0, 0, 1]])
        ctr = V.mean(0)
        ext = V.max(0) - V.min(0)
        return (V - ctr) @ (Rz @ Ry @ Rx).T * scl[k] + ctr + shf[k] * ext

    srv = serve.DeformServer(mem_budget_bytes=args.budget_gb * 1e9)
    # load the kernels (untimed): one tiny request on a target outside the pool
    w = synth.box_union_pair(4, n_boxes=4, n_verts=1000)
    srv.submit(serve.DeformRequest(w[0], w[1], "warm", w[2], w[3], grid=32)).result()
    futs, first = [], set()
    miss = np.zeros(n_req, bool)
    t_start = time.perf_counter()
    t_next = t_start
    for k, (gap, i) in enumerate(zip(gaps, picks)):
        t_next += gap
        now = time.perf_counter()
        if t_next > now:
            time.sleep(t_next - now)
        p = pool[i]
        miss[k] = i not in first
        first.add(i)
        futs.append(srv.submit(serve.DeformRequest(moved(p[0], k), p[1], int(i), p[2], p[3], grid=grids[i])))
        if (k + 1) % int(max(args.rate * 10, 1)) == 0:
            print(f"[bench] serve: {k + 1}/{n_req} submitted, {srv.completed} done", file=sys.stderr, flush=True)
    res = [f.result() for f in futs]
    t_end = time.perf_counter()
    srv.close()
    lat = np.array([r.latency_s for r in res]) * 1e3
    nv = np.array([len(pool[i][0]) for i in picks])
    classes = np.array([c for c, _ in srv.batches])
    sizes_b = np.array([n for _, n in srv.batches])
    pct = lambda x, q: float(np.percentile(x, q)) if len(x) else None   # noqa: E731
    return {"metric": "serving: deformation pairs/sec with p50/p99 latency (config 5, Poisson arrivals, cold cache)",
            "value": len(res) / (t_end - t_start), "unit": "pairs/s", "higher_is_better": True,
            "requests": len(res), "offered_rate": args.rate, "arrival_s": duration,
            "wall_s": t_end - t_start, "dtype": "f32 (LM), f64 (grid)", "data": "synthetic",
            "latency_ms": {"p50": pct(lat, 50), "p99": pct(lat, 99), "max": float(lat.max()),
                           "p50_small_pairs": pct(lat[nv <= 8192], 50), "p99_small_pairs": pct(lat[nv <= 8192], 99),
                           "p50_large_pairs": pct(lat[nv > 8192], 50), "p99_large_pairs": pct(lat[nv > 8192], 99),
                           "p50_cache_miss": pct(lat[miss], 50), "p99_cache_miss": pct(lat[miss], 99)},
            "template_builds": srv.builds, "template_rebuilds_in_place": srv.rebuilds, "evictions": srv.evictions,
            "cache_misses": int(miss.sum()),
            "batches": {"small": int((classes == 0).sum()), "large": int((classes == 1).sum()),
                        "mean_small": float(sizes_b[classes == 0].mean()) if (classes == 0).any() else 0.0,
                        "mean_large": float(sizes_b[classes == 1].mean()) if (classes == 1).any() else 0.0},
            "config": {"workload": "config 5 serving: config-1/config-2-like pairs 50/50, sources log-uniform "
                                   "2k-%d verts, grids 64/128/256^3, fresh rigid motion per request" % args.max_verts,
                       "pool_pairs": n_pool, "grids": {str(g): grids.count(g) for g in (64, 128, 256)},
                       "budget_gb": args.budget_gb, "workers": "3 small + 1 large", "seed": 7}}


def bench_serve5k(args):
    """The headline workload in serving form: Poisson arrivals at args.rate
    requests/s (default 4000) for args.duration seconds (default 10) of
    5000-vertex box-union pairs at 64^3 (128 distinct targets, the template
    cache starting cold; every request moves its pool source by a fresh
    rigid motion) through the continuous-batching server.  Beside it, in the
    same run, the plain one-pair-at-a-time GPU path on the same kind of
    requests: one pair per mo_deform_batch call on the cached template, and
    one pair per rigid_deform pipeline call (normalization and grid per
    call).  Energies are checked against the oracle by the serving and LM
    parity tests; here every result must report convergence or the
    iteration cap."""
    import numpy as np
    import torch
    from meshode_amd import deform, pyDeform, serve, synth
    rate = args.rate if args.rate != 200.0 else 4000.0
    duration = args.duration if args.duration is not None else 10.0
    rng = np.random.default_rng(11)
    pool = [synth.box_union_pair(s) for s in range(128)]
    n_req = int(rate * duration)
    gaps = rng.exponential(1.0 / rate, n_req)
    picks = rng.integers(0, len(pool), n_req)
    ang = np.radians(rng.uniform(-3, 3, (n_req, 3)))
    scl = rng.uniform(0.98, 1.02, n_req)
    shf = rng.uniform(-0.01, 0.01, (n_req, 3))

    def moved(V, k):
        a, b, c = ang[k]
        Rx = np.array([[1, 0, 0], [0, np.cos(a), -np.sin(a)], [0, np.sin(a), np.cos(a)]])
        Ry = np.array([[np.cos(b), 0, np.sin(b)], [0, 1, 0], [-np.sin(b), 0, np.cos(b)]])
        Rz = np.array([[np.cos(c), -np.sin(c), 0], [np.sin(c), np.cos(c), 0], [0, 0, 1]])
        ctr = V.mean(0)
        ext = V.max(0) - V.min(0)
        return (V - ctr) @ (Rz @ Ry @ Rx).T * scl[k] + ctr + shf[k] * ext

    # the requests' sources are generated before the clock starts (the
    # server's own work per request -- admission, upload, solve, download --
    # is what is timed)
    srcs = [moved(pool[i][0], k) for k, i in enumerate(picks)]
    srv = serve.DeformServer(mem_budget_bytes=args.budget_gb * 1e9)
    w = synth.box_union_pair(4, n_boxes=4, n_verts=1000)
    srv.submit(serve.DeformRequest(w[0], w[1], "warm", w[2], w[3], grid=32)).result()
    futs = []
    t_start = time.perf_counter()
    t_next = t_start
    for k, (gap, i) in enumerate(zip(gaps, picks)):
        t_next += gap
        now = time.perf_counter()
        if t_next > now:
            time.sleep(t_next - now)
        p = pool[i]
        futs.append(srv.submit(serve.DeformRequest(srcs[k], p[1], int(i), p[2], p[3], grid=64)))
        if (k + 1) % int(max(rate * 2, 1)) == 0:
            print(f"[bench] serve5k: {k + 1}/{n_req} submitted, {srv.completed} done", file=sys.stderr, flush=True)
    res = [f.result() for f in futs]
    t_end = time.perf_counter()
    srv.close()
    lat = np.array([r.latency_s for r in res]) * 1e3
    terms = collections.Counter(int(r.stats["termination"]) for r in res)
    served = len(res) / (t_end - t_start)
    sizes_b = np.array([n for _, n in srv.batches])
    # plain path 1: one pair per mo_deform_batch call, template cached
    pids = {}
    n_plain = 48
    torch.cuda.synchronize()
    for k in range(4):   # templates and kernels (untimed)
        p = pool[k]
        pids[k] = deform.template_f64(p[2], p[3], 64)
    V0 = []
    for k in range(n_plain):
        p = pool[k % 4]
        _, scale, pos = pyDeform.TemplateInfo(pids[k % 4])
        V0.append(torch.from_numpy((moved(p[0], k) - pos) / scale).cuda())
    deform.deform_batch([pids[0]], [V0[0].clone()], [pool[0][1]], [0], [1.0])
    torch.cuda.synchronize()
    tp = time.perf_counter()
    for k in range(n_plain):
        deform.deform_batch([pids[k % 4]], [V0[k]], [pool[k % 4][1]], [0], [1.0])
    torch.cuda.synchronize()
    plain_cached = n_plain / (time.perf_counter() - tp)
    # plain path 2: one pair per rigid_deform pipeline call (grid built per call)
    n_pipe = 16
    deform.rigid_deform_batch([pool[0]], grid_resolution=64)
    torch.cuda.synchronize()
    tp = time.perf_counter()
    for k in range(n_pipe):
        deform.rigid_deform_batch([pool[k]], grid_resolution=64)
    torch.cuda.synchronize()
    plain_pipe = n_pipe / (time.perf_counter() - tp)
    pct = lambda x, q: float(np.percentile(x, q))   # noqa: E731
    return {"metric": "serving: deformation pairs/sec with p50/p99 latency (5k verts, 64^3, Poisson arrivals)",
            "value": served, "unit": "pairs/s", "higher_is_better": True,
            "requests": len(res), "offered_rate": rate, "arrival_s": duration, "wall_s": t_end - t_start,
            "dtype": "f32 (LM), f64 (grid)", "data": "synthetic",
            "latency_ms": {"p50": pct(lat, 50), "p99": pct(lat, 99), "max": float(lat.max())},
            "terminations": {str(k): v for k, v in sorted(terms.items())},
            "plain_one_pair_per_call_cached_template_pairs_per_s": plain_cached,
            "plain_one_pair_rigid_deform_pipeline_pairs_per_s": plain_pipe,
            "served_over_plain_cached": served / plain_cached, "served_over_plain_pipeline": served / plain_pipe,
            "template_builds": srv.builds, "batches": int(len(sizes_b)),
            "mean_batch": float(sizes_b.mean()) if len(sizes_b) else 0.0,
            "config": {"workload": "headline pairs served: 12-box unions of 5000 verts vs jittered targets, 64^3, "
                                   "128 distinct targets (cold cache), fresh rigid motion per request",
                       "budget_gb": args.budget_gb, "workers": "3 small + 1 large", "seed": 11}}


def bench_adam(args):
    """Config 2: RigidLossLayer + Adam (lr 1e-3) on a 5000-vertex box union vs
    its target at 128^3; steps/s of the forward/backward/Adam loop."""
    import numpy as np
    import torch
    from meshode_amd import layers, synth
    sV, sF, tV, tF = synth.box_union_pair(0, n_verts=5000)
    src = torch.from_numpy(sV.astype(np.float32)).cuda()
    F = torch.from_numpy(sF).cuda()
    layer = layers.RigidLossLayer(src, F, torch.from_numpy(tV.astype(np.float32)).cuda(),
                                  torch.from_numpy(tF).cuda(), grid_resolution=128)
    V = torch.nn.Parameter(src.clone())
    opt = torch.optim.Adam([V], lr=1e-3)
    steps = max(args.steps, 1) * 100
    for _ in range(20):
        opt.zero_grad(); layer(V, F).backward(); opt.step()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        opt.zero_grad()
        loss = layer(V, F)
        loss.backward()
        opt.step()
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    return {"metric": "pyDeform Adam steps/sec (5000 verts, 128^3, rigidity 1)", "value": steps / dt,
            "unit": "steps/s", "steps": steps, "final_loss": float(loss.item()),
            "config": {"workload": "config 2: RigidLossLayer + Adam lr 1e-3", "verts": 5000, "grid": 128}}


def bench_ode(args):
    """Config 3: NeuralODE flow (3-256-256-256-3 softplus, RK4 x 16, adjoint)
    pushing 50k surface points + the 5000 mesh vertices, loss on a 128^3 grid
    + rigidity 0.1; optimizer steps/s."""
    import numpy as np
    import torch
    from meshode_amd import neuralode, pyDeform, synth
    sV, sF, tV, tF = synth.box_union_pair(0, n_verts=5000)
    pid = pyDeform.InitializeDeformTemplate(torch.from_numpy(tV.astype(np.float32)), torch.from_numpy(tF), 0, 128)
    V = torch.from_numpy(sV.astype(np.float32))
    pyDeform.NormalizeByTemplate(V, pid)
    rng = np.random.default_rng(0)
    f = rng.integers(0, len(sF), 50000)
    w = rng.dirichlet([1, 1, 1], 50000)
    pts = (V.numpy()[sF[f]] * w[:, :, None]).sum(1).astype(np.float32)      # uniform-in-face samples
    P = torch.cat([V, torch.from_numpy(pts)]).cuda()
    E = torch.from_numpy(np.concatenate([sF[:, [0, 1]], sF[:, [1, 2]], sF[:, [2, 0]]]).astype(np.int32)).cuda()
    model = neuralode.NeuralODEDeformer(P, E, pid, n_mesh=len(V), rho=0.1, steps=16)
    opt = torch.optim.Adam(model.field.parameters(), lr=1e-3)
    for _ in range(2):
        opt.zero_grad(); model()[0].backward(); opt.step()
    torch.cuda.synchronize()
    steps = max(args.steps, 1) * 5
    t0 = time.perf_counter()
    for _ in range(steps):
        opt.zero_grad()
        loss, _ = model()
        loss.backward()
        opt.step()
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    flops = 2 * P.shape[0] * (3 * 256 + 2 * 256 * 256 + 256 * 3) * 64 * 4     # fwd + recompute + 2x bwd
    return {"metric": "NeuralODE optimizer steps/sec (55k points, RK4x16, 128^3)", "value": steps / dt,
            "unit": "steps/s", "steps": steps, "final_loss": float(loss.item()),
            "approx_mlp_tflops": flops * steps / dt / 1e12,
            "config": {"workload": "config 3: velocity MLP 3-256-256-256-3 softplus, RK4x16, adjoint", "points":
                       int(P.shape[0]), "grid": 128}}


def bench_flow256(args):
    """Config 3's forward integration alone: 16 fused RK4 steps
    (mo_flow256_step, FP32 MFMA) of the 3-256-256-256-3 softplus field over
    55k points; kernel time from HIP events on the launch stream, MFMA
    roofline at the FP32 dense peak."""
    import torch
    from meshode_amd import neuralode
    n, steps = 55000, 16
    torch.manual_seed(0)
    y = torch.rand(n, 3, device="cuda")
    field = neuralode.VelocityMLP(256, seed=0).cuda()
    packed = neuralode.flow256_pack(field)
    for _ in range(2):
        yy = y
        for _ in range(steps):
            yy = neuralode.flow256_step(packed, yy, 1.0 / steps)
    torch.cuda.synchronize()
    reps = max(args.steps, 1) * 5
    stream = torch.cuda.current_stream()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    t0 = time.perf_counter()
    e0.record(stream)
    for _ in range(reps):
        yy = y
        for _ in range(steps):
            yy = neuralode.flow256_step(packed, yy, 1.0 / steps)
    e1.record(stream)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    ms_step = e0.elapsed_time(e1) / (reps * steps)
    flops = 2.0 * n * (3 * 256 + 2 * 256 * 256 + 256 * 3) * 4      # four field evaluations per step
    tflops = flops / (ms_step * 1e-3) / 1e12
    return {"metric": "config-3 forward RK4 steps/sec (55k points, fused FP32 MFMA step)",
            "value": reps * steps / wall, "unit": "steps/s", "steps": reps * steps, "ms_per_step": ms_step,
            "higher_is_better": True, "dtype": "f32", "data": "synthetic",
            "roofline": {"bound": "mfma", "achieved": tflops, "peak": 157.3, "unit": "TFLOP/s",
                         "frac": tflops / 157.3, "traffic": None},
            "config": {"workload": "velocity MLP 3-256-256-256-3 softplus, one RK4 step per launch", "points": n,
                       "flops_per_step": flops}}


def bench_rot(args):
    """rigid_rot_deform's solve (Deformer::DeformWithRot) over a batch: the
    headline's pairs (args.pairs, args.distinct distinct seeded 12-box unions
    of args.verts vertices vs jittered targets, 64^3, lambda 1), templates
    built once outside the timed region; solves/s of one
    mo_deform_rot_batch call from the rest shapes.  Also the one-pair
    latency (mo_deform_rot).  The oracle's exact-step LM on pair 0 gives the
    energy check and, with its grid build, the CPU baseline (one thread)."""
    import numpy as np
    import torch
    from meshode_amd import deform, pyDeform, synth
    distinct = [synth.box_union_pair(s, n_verts=args.verts) for s in range(min(args.distinct, args.pairs))]
    pids, rest, faces = [], [], []
    for sV, sF, tV, tF in distinct:
        pid = deform.template_f64(tV, tF, args.grid)
        _, scale, trans = pyDeform.TemplateInfo(pid)
        pids.append(pid)
        rest.append(torch.from_numpy((sV - trans) / scale).cuda().contiguous())
        faces.append(torch.from_numpy(sF).cuda())
    idx = [i % len(distinct) for i in range(args.pairs)]
    Vs = [rest[i].clone() for i in idx]
    # one-pair latency
    V1 = rest[0].clone()
    deform.deform_rot(pids[0], V1, faces[0], args.lam)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    V1.copy_(rest[0])
    st1, W1 = deform.deform_rot(pids[0], V1, faces[0], args.lam)
    torch.cuda.synchronize()
    single_ms = 1e3 * (time.perf_counter() - t0)
    print(f"[bench] one pair: {single_ms:.1f} ms", file=sys.stderr, flush=True)
    # the same pair on a team of workgroups (64-thread workgroups, 512 vertices each)
    topt = deform.default_options(pcg_relative_tolerance=1e-6, block_threads=64, team_vertices=512)
    V1.copy_(rest[0])
    deform.deform_rot(pids[0], V1, faces[0], args.lam, options=topt)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    V1.copy_(rest[0])
    stt, _ = deform.deform_rot(pids[0], V1, faces[0], args.lam, options=topt)
    torch.cuda.synchronize()
    team_ms = 1e3 * (time.perf_counter() - t0)
    print(f"[bench] one pair on a team: {team_ms:.1f} ms", file=sys.stderr, flush=True)
    for _ in range(max(args.warmup, 0)):
        deform.deform_rot_batch([pids[i] for i in idx], Vs, [faces[i] for i in idx], [args.lam] * args.pairs)
        for V, i in zip(Vs, idx):
            V.copy_(rest[i])
    torch.cuda.synchronize()
    steps = max(args.steps, 1)
    elapsed = 0.0
    for k in range(steps):
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        stb, _ = deform.deform_rot_batch([pids[i] for i in idx], Vs, [faces[i] for i in idx], [args.lam] * args.pairs)
        torch.cuda.synchronize()
        dt = time.perf_counter() - t0
        elapsed += dt
        print(f"[bench] step {k + 1}/{steps}: {1e3 * dt:.1f} ms", file=sys.stderr, flush=True)
        for V, i in zip(Vs, idx):
            V.copy_(rest[i])
    value = args.pairs * steps / elapsed
    out = {"metric": "rigid_rot_deform solves/sec (5k verts, 64^3, energy tol 1e-3)", "value": value,
           "unit": "solves/s", "n_gpus": 1, "steps": steps, "warmup": args.warmup,
           "ms_per_step": 1e3 * elapsed / steps, "higher_is_better": True,
           "dtype": "f32 (LM vectors), f64 (reductions)", "data": "synthetic",
           "single_pair_ms": single_ms,
           "single_pair_team_ms": team_ms,
           "gpu_stats_pair0": {k: st1[k] for k in ("final_cost", "iterations", "pcg_iterations", "termination")},
           "gpu_stats_pair0_team": {k: stt[k] for k in ("final_cost", "iterations", "pcg_iterations", "termination")},
           "max_iterations_in_batch": int(max(s["iterations"] for s in stb)),
           "config": {"workload": "rigid_rot_deform: %d pairs (%d distinct 12-box unions of %d verts vs jittered "
                                  "targets)" % (args.pairs, len(distinct), args.verts),
                      "grid": args.grid, "lambda": args.lam}}
    if not args.no_cpu_baseline:
        from oracle import oracle as O
        O.build()
        sV, sF, tV, tF = distinct[0]
        t0 = time.perf_counter()
        tn, sc, pos = O.normalize(tV)
        grid, _, _ = O.build_grid_bvh(tn, tF, args.grid)
        sno = O.apply_transform(sV, sc, pos)
        _, _, so = O.lm_rot_solve(grid, args.grid, sno, sF, args.lam)
        tc = time.perf_counter() - t0
        g64 = pyDeform.TemplateGrid(pids[0])[0].cpu().numpy().reshape(-1)
        eg = O.lm_rot_cost(g64, args.grid, V1.cpu().numpy(), W1.cpu().numpy(), rest[0].cpu().numpy(), sF,
                           args.lam)[0]
        out["cpu_baseline"] = {"value": 1.0 / tc, "unit": "solves/s", "cores": 1, "kind": "port",
                               "sample": "one full solve of pair 0 (grid build + exact-step LM, FP64)"}
        out["energy_check_pair0"] = {"gpu": eg, "oracle": so["final_cost"],
                                     "rel_diff": abs(eg - so["final_cost"]) / so["final_cost"]}
    return out


def _neural2_meshes():
    """Source and target CAD meshes: 12-box unions of 2000 input vertices
    each, through LoadCadMesh (subdivision 2e-2, graph cells 1e-2)."""
    import os
    import tempfile

    from meshode_amd import objio, pyDeform, synth
    sV, sF, tV, tF = synth.box_union_pair(0, n_verts=2000)
    d = tempfile.mkdtemp()
    out = []
    for name, V, F in (("source", sV, sF), ("target", tV, tF)):
        p = os.path.join(d, name + ".obj")
        objio.write_obj(p, V, F)
        out.append(pyDeform.LoadCadMesh(p))
    return out


def bench_neural2(args):
    """NeuralODE bidirectional CAD deformation (cad_neural_deform2.py):
    iterations/s of the Adam loop (flow forward on graph 1, inverse on graph
    2, GraphLoss2 both ways, ReverseLoss both ways, backward, Adam), plus the
    final LinearEstimationWithRot on the source mesh."""
    import numpy as np
    import torch
    from meshode_amd.cad_neural_deform2 import NeuralDeform2
    m1, m2 = _neural2_meshes()
    raw = [m[4].clone() for m in (m1, m2)]
    torch.manual_seed(0)
    d = NeuralDeform2(m1, m2, 0.1, torch.device("cuda"))
    for _ in range(5):
        d.step()
    torch.cuda.synchronize()
    iters = max(args.steps, 1) * 40
    t0 = time.perf_counter()
    rows = [d.step() for _ in range(iters)]
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    last = torch.stack(rows[-1:]).cpu().numpy()[0]
    t1 = time.perf_counter()
    V = d.finalize()
    torch.cuda.synchronize()
    t_fin = time.perf_counter() - t1
    out = {"metric": "cad_neural_deform2 Adam iterations/sec", "value": iters / dt, "unit": "iterations/s",
           "n_gpus": 1, "steps": iters, "warmup": 5, "ms_per_step": 1e3 * dt / iters, "higher_is_better": True,
           "dtype": "f32 (flow, losses); f64 (nearest distances, LinearEstimationWithRot)", "data": "synthetic",
           "final_losses": [float(x) for x in last], "finalize_ms": 1e3 * t_fin,
           "finite_output": bool(np.isfinite(V.numpy()).all()),
           "config": {"workload": "cad_neural_deform2: 12-box unions (2000 input verts) through LoadCadMesh",
                      "source_verts": int(m1[0].shape[0]), "graph_verts": [int(m1[4].shape[0]), int(m2[4].shape[0])],
                      "graph_edges": [int(m1[5].shape[0]), int(m2[5].shape[0])], "grid": 64, "rigidity": 0.1}}
    if not args.no_cpu_baseline:
        out["cpu_baseline"] = _neural2_cpu_baseline(m1, m2, raw, d)
    return out


def _neural2_cpu_baseline(m1, m2, raw, d, iters=3):
    """The same loop as the reference runs it on the host: torch CPU flow and
    Adam, the oracle's C restatement of the pyDeform layers, scipy cKDTree
    for ReverseLossLayer (as reverse_loss_layer.py:18-19 does)."""
    import numpy as np
    import torch
    from scipy.spatial import cKDTree
    from torch.autograd import Function

    from meshode_amd import pyDeform
    from meshode_amd.layers.neuralode_fast import NeuralODE
    from oracle import oracle as O
    O.build()
    pid1, pid2 = int(d.param_id1), int(d.param_id2)
    info = [pyDeform.TemplateInfo(p) for p in (pid1, pid2)]
    grids = [pyDeform.TemplateGrid(p)[0].cpu().numpy().reshape(-1) for p in (pid1, pid2)]
    G = [torch.from_numpy(O.normalize_by_template_f32(raw[k].numpy(), info[k][1], info[k][2])) for k in (0, 1)]
    E = [m1[5].numpy(), m2[5].numpy()]
    off = [O.store_graph(G[k].numpy(), E[k]) for k in (0, 1)]
    r2 = torch.tensor(0.1 * 0.1).tolist()

    class GL2(Function):
        @staticmethod
        def forward(ctx, V, k_grid, k_edges):
            v = V.detach().numpy()
            N = info[k_grid][0]
            D = O.distance_loss_fwd(grids[k_grid], N, v) * np.float32(0.5)
            R = O.graph_fwd(v, E[k_edges], off[k_edges]) * np.float32(0.5)
            ctx.save_for_backward(V)
            ctx.k = (k_grid, k_edges)
            return torch.from_numpy(D).sum() + torch.from_numpy(R).sum() * r2

        @staticmethod
        def backward(ctx, g):
            (V,) = ctx.saved_tensors
            k_grid, k_edges = ctx.k
            v = V.detach().numpy()
            gD = torch.from_numpy(O.distance_loss_bwd(grids[k_grid], info[k_grid][0], v))
            gR = torch.from_numpy(O.graph_bwd(v, E[k_edges], off[k_edges]))
            return g * (gD + gR * r2), None, None

    def rev(src, tar):
        _, ii = cKDTree(src.detach().numpy()).query(tar.numpy(), k=1, workers=4)
        x = src[torch.from_numpy(ii)] - tar
        return 0.5 * (x * x).sum()

    torch.manual_seed(0)
    func = NeuralODE(torch.device("cpu"))
    opt = torch.optim.Adam(func.parameters(), lr=1e-3)
    t0 = time.perf_counter()
    for _ in range(iters):
        opt.zero_grad()
        a = func.forward(G[0])
        b = func.inverse(G[1])
        loss = GL2.apply(a, 1, 0) + rev(a, G[1]) + GL2.apply(b, 0, 1) + rev(b, G[0])
        loss.backward()
        opt.step()
    dt = time.perf_counter() - t0
    return {"value": iters / dt, "unit": "iterations/s", "cores": torch.get_num_threads(), "kind": "port",
            "sample": "%d iterations of the same loop on the host (torch CPU flow, oracle layers, cKDTree with 4 "
                      "workers)" % iters}


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
