"""Autotuner for the launch shapes of the batched LM solver.

The register/LDS kernel ``k_lm_solve_lds<T, VPT>`` has one launch shape per
(threads per pair T, vertices per thread VPT).  A pair of nv vertices can
run on any shape with T * VPT >= nv, and which one is fastest depends on nv
and on the mesh: a larger VPT holds more per-vertex vectors in registers
(some shapes spill), a larger T leaves more lanes idle on small pairs and
needs more waves per barrier.  The tuner times every fitting shape on the
caller's own pairs, one size class at a time, with the library's phase
timer (HIP events around the LM launch), and installs the fastest shape of
each class as the library's size-class table (``mo_lm_set_shape_table``).
Pairs that the register/LDS kernel cannot take (more than 8192 vertices, or
edge multiplicities above 2) run on the global-memory kernels, a team of
ceil(nv / team_vertices) co-resident workgroups of ``block_threads``
threads per pair; those two options are tuned the same way and returned
for the caller's options.

The reference has no counterpart (its Ceres solve is single-threaded,
deformer.cc:55-59).  A shape changes only how a pair's vertices are spread
over threads, so every choice is the same algorithm and meets the same
energy bar against the oracle; results are bit-identical between calls with
the same shape.
"""
import ctypes as C
import statistics

import torch

from . import _lib, deform

LDS_MAX_VERTICES = 8192


def shapes():
    """[(threads, vertices_per_thread)] of the register/LDS kernel, by index."""
    lib = _lib.load()
    out = []
    for i in range(lib.mo_lm_shape_count()):
        t, v = C.c_int32(), C.c_int32()
        _lib.call("mo_lm_shape_info", i, C.byref(t), C.byref(v))
        out.append((t.value, v.value))
    return out


def set_table(rows):
    """Install a size-class table [(max_vertices, shape_index)] (ascending
    max_vertices); an empty list restores the built-in order."""
    n = len(rows)
    mx = (C.c_int32 * max(n, 1))(*[int(r[0]) for r in rows])
    sh = (C.c_int32 * max(n, 1))(*[int(r[1]) for r in rows])
    _lib.call("mo_lm_set_shape_table", C.cast(mx, C.c_void_p) if n else None,
              C.cast(sh, C.c_void_p) if n else None, n)


def size_classes():
    """Upper bounds of the size classes: the distinct capacities T * VPT."""
    return sorted({t * v for t, v in shapes()})


def _lm_ms(batch, options, repeats):
    batch.run(options)                 # warm-up: loads the shape's code object
    times = []
    for _ in range(repeats):
        batch.run(options)
        times.append(_lib.last_timings()["lm_ms"])
    return float(statistics.median(times))


def tune(pairs, grid_resolution=64, lambdas=1.0, options=None, repeats=3, max_pairs=256,
         block_threads=(512, 1024), team_factors=(2, 4, 8), install=True):
    """Time the launch shapes on ``pairs`` [(src_V, src_F, tgt_V, tgt_F)] as
    the rigid_deform pipeline runs them (``deform.PairBatch``) and pick the
    fastest per size class.

    grid_resolution / lambdas: one value or one per pair; options: the
    solver options of the real run (iteration caps and tolerances change
    the work, so tune with the ones you will use); at most ``max_pairs``
    pairs per class are timed.  With ``install`` the table is installed for
    every later call in this process (else the built-in order is restored).
    Pairs above 8192 vertices are timed at every block_threads and
    team_vertices = factor * block_threads.
    Returns {"table": [(max_vertices, threads, vpt, lm_ms)], "candidates":
    {max_vertices: {"TxVPT": lm_ms}}, "block_threads" and "team_vertices":
    the best (None without large pairs), "large_ms": {"bt/tv": lm_ms}}."""
    n = len(pairs)
    gr = list(grid_resolution) if isinstance(grid_resolution, (list, tuple)) else [grid_resolution] * n
    lam = list(lambdas) if isinstance(lambdas, (list, tuple)) else [lambdas] * n
    base = options or deform.default_options()
    opt = deform.default_options()
    C.memmove(C.byref(opt), C.byref(base), C.sizeof(opt))
    opt.timing = 1
    shp = shapes()
    bounds = size_classes()
    classes = {}
    large = []
    for i, p in enumerate(pairs):
        nv = len(p[0])
        if nv > LDS_MAX_VERTICES:
            large.append(i)
            continue
        ub = next(b for b in bounds if b >= nv)
        classes.setdefault(ub, []).append(i)

    def sub(idx):
        idx = idx[:max_pairs]
        return deform.PairBatch([pairs[i] for i in idx], grid_resolution=[gr[i] for i in idx],
                                lambdas=[lam[i] for i in idx])

    table, cand_ms = [], {}
    try:
        for ub in sorted(classes):
            batch = sub(classes[ub])
            ms = {}
            for b, (t, v) in enumerate(shp):
                if t * v < ub:
                    continue
                set_table([(ub, b)])
                ms[b] = _lm_ms(batch, opt, repeats)
            best = min(ms, key=ms.get)
            table.append((ub, best))
            cand_ms[ub] = {f"{shp[b][0]}x{shp[b][1]}": m for b, m in ms.items()}
            del batch
    finally:
        set_table(table if install else [])
    large_ms, best = {}, (None, None)
    if large:
        batch = sub(large)
        for bt in block_threads:
            for f in team_factors:
                opt.block_threads, opt.team_vertices = int(bt), int(f * bt)
                large_ms[(int(bt), int(f * bt))] = _lm_ms(batch, opt, repeats)
        best = min(large_ms, key=large_ms.get)
        del batch
    torch.cuda.synchronize()
    return {"table": [(ub, shp[b][0], shp[b][1], cand_ms[ub][f"{shp[b][0]}x{shp[b][1]}"]) for ub, b in table],
            "candidates": cand_ms, "block_threads": best[0], "team_vertices": best[1],
            "large_ms": {f"{bt}/{tv}": m for (bt, tv), m in large_ms.items()}}


def describe(result):
    """One-line summary of a tune() result, e.g.
    '<=5120: 1024x5; teams: 1024 threads, 4096 vertices per workgroup'."""
    parts = [f"<={ub}: {t}x{v}" for ub, t, v, _ in result["table"]]
    if result["block_threads"]:
        parts.append(f"teams: {result['block_threads']} threads, {result['team_vertices']} vertices per workgroup")
    return "; ".join(parts) if parts else "built-in"
