"""Seeded synthetic source/target mesh pairs for the deformation workloads.

The reference ships only four sample meshes and no generator; these
workloads are defined here so that every test and benchmark is reproducible
from a seed (numpy PCG64):

* ``icosphere(level)``: subdivided icosahedron on the unit sphere,
  V = 10 * 4**level + 2, F = 20 * 4**level.
* ``displaced_sphere_pair(seed)``: source = icosphere level 4 (radius 1);
  target = icosphere level 5 with radius
  r = 1 + 0.15 * sum_{k=1..4} a_k sin(f_k theta + phi_k) cos(g_k varphi)
  (theta = atan2(y, x), varphi = arccos(z), a_k ~ U(0.3, 1),
  f_k, g_k in {1..4}, phi_k ~ U(0, 2 pi)), then scaled by (1.0, 0.7, 1.3).
* ``box_union_pair(seed, n_verts=5000)``: source = 12 seeded axis-aligned
  boxes (extents U(0.05, 0.4), centres U(-0.5, 0.5)), each a closed
  2-manifold grid-tessellated box surface; the components overlap and are
  not merged (non-manifold as a union, disconnected), with the per-axis
  segment counts chosen so that the total is exactly ``n_verts`` vertices.
  Target = the same boxes with extents jittered x U(0.8, 1.25) and centres
  offset by N(0, 0.03).
"""
from __future__ import annotations

import numpy as np


def icosphere(level: int):
    t = (1.0 + 5.0 ** 0.5) / 2.0
    V = [(-1, t, 0), (1, t, 0), (-1, -t, 0), (1, -t, 0), (0, -1, t), (0, 1, t),
         (0, -1, -t), (0, 1, -t), (t, 0, -1), (t, 0, 1), (-t, 0, -1), (-t, 0, 1)]
    V = [np.array(v, float) / np.linalg.norm(v) for v in V]
    F = [(0, 11, 5), (0, 5, 1), (0, 1, 7), (0, 7, 10), (0, 10, 11), (1, 5, 9), (5, 11, 4),
         (11, 10, 2), (10, 7, 6), (7, 1, 8), (3, 9, 4), (3, 4, 2), (3, 2, 6), (3, 6, 8),
         (3, 8, 9), (4, 9, 5), (2, 4, 11), (6, 2, 10), (8, 6, 7), (9, 8, 1)]
    for _ in range(level):
        cache = {}

        def mid(a, b):
            key = (a, b) if a < b else (b, a)
            if key not in cache:
                m = V[a] + V[b]
                V.append(m / np.linalg.norm(m))
                cache[key] = len(V) - 1
            return cache[key]

        nf = []
        for a, b, c in F:
            ab, bc, ca = mid(a, b), mid(b, c), mid(c, a)
            nf += [(a, ab, ca), (b, bc, ab), (c, ca, bc), (ab, bc, ca)]
        F = nf
    return np.array(V, np.float64), np.array(F, np.int32)


def displaced_sphere_pair(seed: int, src_level: int = 4, tgt_level: int = 5):
    rng = np.random.Generator(np.random.PCG64(seed))
    a = rng.uniform(0.3, 1.0, 4)
    f = rng.integers(1, 5, 4)
    g = rng.integers(1, 5, 4)
    ph = rng.uniform(0.0, 2 * np.pi, 4)
    sV, sF = icosphere(src_level)
    tV, tF = icosphere(tgt_level)
    theta = np.arctan2(tV[:, 1], tV[:, 0])
    varphi = np.arccos(np.clip(tV[:, 2], -1.0, 1.0))
    r = 1.0 + 0.15 * sum(a[k] * np.sin(f[k] * theta + ph[k]) * np.cos(g[k] * varphi) for k in range(4))
    tV = tV * r[:, None] * np.array([1.0, 0.7, 1.3])
    return sV, sF, tV, tF


def box_surface(center, extent, segs):
    """Closed, outward-oriented triangulated surface of an axis-aligned box
    with segs = (nx, ny, nz) segments per axis."""
    nx, ny, nz = (int(s) for s in segs)
    n = np.array([nx, ny, nz])
    lo = np.asarray(center, float) - 0.5 * np.asarray(extent, float)
    step = np.asarray(extent, float) / n
    I, J, K = np.meshgrid(np.arange(nx + 1), np.arange(ny + 1), np.arange(nz + 1), indexing="ij")
    on = (I == 0) | (I == nx) | (J == 0) | (J == ny) | (K == 0) | (K == nz)
    idmap = -np.ones(on.shape, np.int64)
    # vertex order: lattice order (i, j, k) restricted to the surface
    idmap[on] = np.arange(int(on.sum()))
    verts = lo + step * np.stack([I[on], J[on], K[on]], 1).astype(float)
    faces = []
    for axis in range(3):
        u, v = (axis + 1) % 3, (axis + 2) % 3
        A, B = np.meshgrid(np.arange(n[u]), np.arange(n[v]), indexing="ij")
        A, B = A.reshape(-1), B.reshape(-1)
        for side in (0, 1):
            def P(da, db):
                c = [None, None, None]
                c[axis] = np.full_like(A, side * n[axis])
                c[u] = A + da
                c[v] = B + db
                return idmap[c[0], c[1], c[2]]
            p00, p10, p11, p01 = P(0, 0), P(1, 0), P(1, 1), P(0, 1)
            if side == 1:
                t = np.stack([np.stack([p00, p10, p11], 1), np.stack([p00, p11, p01], 1)], 1)
            else:
                t = np.stack([np.stack([p00, p11, p10], 1), np.stack([p00, p01, p11], 1)], 1)
            faces.append(t.reshape(-1, 3))
    return verts, np.concatenate(faces).astype(np.int32)


def _box_verts(s):
    nx, ny, nz = s
    return 2 * (nx * ny + ny * nz + nz * nx) + 2


def _choose_segments(extents, n_verts):
    """Per-box (nx, ny, nz) segment counts, proportional to the extents, whose
    closed surfaces total exactly n_verts vertices (one box is re-solved
    exactly; the density is nudged until a solution exists)."""
    ext = np.asarray(extents)

    def total(segs):
        return sum(_box_verts(s) for s in segs)

    lo, hi = 0.1, 1000.0
    for _ in range(100):
        rho = 0.5 * (lo + hi)
        segs = [tuple(max(1, int(round(rho * e))) for e in ex) for ex in ext]
        if total(segs) > n_verts:
            hi = rho
        else:
            lo = rho
    for nudge in range(200):
        rho = lo * (1.0 - 0.002 * nudge)
        base = [tuple(max(1, int(round(rho * e))) for e in ex) for ex in ext]
        for b in range(len(base) - 1, -1, -1):
            rest = n_verts - (total(base) - _box_verts(base[b]))
            tgt = rest - 2
            if tgt <= 0 or tgt % 2:
                continue
            tgt //= 2
            want = ext[b] * rho
            best = None
            for nx in range(1, 120):
                for ny in range(1, 120):
                    num = tgt - nx * ny
                    if num <= 0:
                        break
                    if num % (nx + ny):
                        continue
                    nz = num // (nx + ny)
                    dev = float(np.sum((np.array([nx, ny, nz]) - want) ** 2))
                    if best is None or dev < best[0]:
                        best = (dev, (nx, ny, nz))
            if best is not None:
                segs = list(base)
                segs[b] = best[1]
                assert total(segs) == n_verts
                return segs
    raise ValueError("cannot reach exact vertex count %d" % n_verts)


def box_union_pair(seed: int, n_boxes: int = 12, n_verts: int = 5000):
    rng = np.random.Generator(np.random.PCG64(seed))
    ext = rng.uniform(0.05, 0.4, (n_boxes, 3))
    cen = rng.uniform(-0.5, 0.5, (n_boxes, 3))
    jit = rng.uniform(0.8, 1.25, (n_boxes, 3))
    off = rng.normal(0.0, 0.03, (n_boxes, 3))
    segs = _choose_segments(ext, n_verts)

    def assemble(centers, extents):
        Vs, Fs, base = [], [], 0
        for c, e, s in zip(centers, extents, segs):
            v, f = box_surface(c, e, s)
            Vs.append(v)
            Fs.append(f + base)
            base += len(v)
        return np.concatenate(Vs), np.concatenate(Fs).astype(np.int32)

    sV, sF = assemble(cen, ext)
    tV, tF = assemble(cen + off, ext * jit)
    return sV, sF, tV, tF


def icosphere_pair(seed: int, level: int):
    """Displaced-sphere pair with source level ``level`` and target level+1."""
    return displaced_sphere_pair(seed, src_level=level, tgt_level=min(level + 1, 6))
