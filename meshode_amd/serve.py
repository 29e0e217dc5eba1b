"""Continuous-batching deformation server on one GPU.

Requests (source mesh + target mesh or a cached target key, grid resolution,
rigidity) arrive at any time.  They are admitted into solve slots as slots
and memory free up:

* Two size classes, each with its own queue and worker threads (one HIP
  stream per worker): pairs that fit the register/LDS solver (at most
  ``team_threshold`` source vertices: one workgroup, one CU, per pair) and
  larger pairs (multi-workgroup teams).  A 100k-vertex pair therefore never
  shares a launch with, or sets the latency of, the small pairs.
* A worker that is free takes every waiting request of its class that fits
  the memory budget, up to ``max_batch`` (256 small pairs = one workgroup per
  CU), and launches them as one batched Levenberg-Marquardt solve
  (Deformer::Deform, deformer.cc:18-92).  With several workers per class,
  requests that arrive while launches are on the GPU are admitted by the
  next worker to free up, and their launches overlap on the device.
* Target templates (normalization + distance grid) stay in HBM in an LRU
  cache keyed by the request's target key.  A missing template is built by
  the worker that needs it, outside the server lock, with a per-key marker
  so concurrent requests for the same target wait for that one build; its
  bytes are charged once, when the build is reserved.  When the budget is
  full, the least recently used idle template of the same grid resolution
  is rebuilt in place for the new target (mo_template_rebuild_f64: no
  device allocation or free, so no device-wide synchronization while other
  launches run); only when none exists are idle templates freed, least
  recently used first, outside the lock.

A request's result does not depend on which other requests share its batch
(each pair is solved by its own workgroup or team with a fixed reduction
order): tests/test_gpu_serve.py checks it bit for bit.
"""
from __future__ import annotations

import collections
import threading
import time
from concurrent.futures import Future
from dataclasses import dataclass, field

import numpy as np
import torch

from . import _lib, deform


@dataclass
class DeformRequest:
    src_V: np.ndarray            # [n,3] FP64, raw coordinates
    src_F: np.ndarray            # [m,3] int32
    tgt_key: object              # cache key of the target
    tgt_V: np.ndarray | None     # [k,3] FP64 (needed unless the key is cached)
    tgt_F: np.ndarray | None
    grid: int = 64
    lam: float = 1.0
    arrival: float = field(default_factory=time.perf_counter)


@dataclass
class DeformResult:
    V: np.ndarray                # deformed source in the target's original frame
    stats: dict
    latency_s: float
    batch_size: int
    template_built: bool = False  # this request's batch built its target's template


def solver_bytes(nv, nf):
    """Device working set of one pair in the batched solver (bytes)."""
    return nv * (2 * 16 + 8 + 4 + 13 * 16 + 24) + 3 * nf * 2 * 4 * 2


def template_bytes(N, nv_t=0, nf_t=0):
    """Device bytes a template holds: FP64 and FP32 grids and the nearest
    triangle index (the target mesh itself is not kept)."""
    return N ** 3 * (8 + 4 + 4)


class _Template:
    __slots__ = ("pid", "N", "scale", "pos", "nbytes", "users", "ready", "error")

    def __init__(self, N, nbytes, pid=-1):
        self.pid, self.N, self.scale, self.pos, self.nbytes, self.users = pid, N, 1.0, None, nbytes, 0
        self.ready = threading.Event()
        self.error = None


class DeformServer:
    def __init__(self, mem_budget_bytes=200e9, max_batch=256, small_workers=3, large_workers=1, options=None,
                 team_threshold=8192, workers=None):
        if not torch.cuda.is_available():
            raise RuntimeError("DeformServer needs a ROCm GPU")
        _lib.load()
        if workers is not None:                       # one pool size for both classes
            small_workers = max(1, int(workers))
        self.budget = float(mem_budget_bytes)
        self.max_batch = int(max_batch)
        self.team_threshold = int(team_threshold)
        self.options = options or deform.default_options()
        self._cv = threading.Condition()
        self._queues = (collections.deque(), collections.deque())     # small, large
        self._cache = collections.OrderedDict()       # key -> _Template (LRU order)
        self._cache_bytes = 0.0                       # resident and in-build templates
        self._inflight_bytes = 0.0                    # solver working sets of admitted requests
        self._closed = False
        self.completed = 0
        self.batches = []                             # (size class, batch size)
        self.builds = 0
        self.rebuilds = 0                             # builds into a recycled template
        self.evictions = 0
        self._doomed = []                             # evicted pids, freed outside the lock
        self._threads = [threading.Thread(target=self._worker, args=(c,), daemon=True)
                         for c, n in ((0, small_workers), (1, large_workers)) for _ in range(max(1, n))]
        for t in self._threads:
            t.start()

    # -- public ----------------------------------------------------------
    def submit(self, req: DeformRequest) -> Future:
        fut = Future()
        cls = 0 if len(req.src_V) <= self.team_threshold else 1
        with self._cv:
            if self._closed:
                raise RuntimeError("server closed")
            self._queues[cls].append((req, fut))
            self._cv.notify_all()
        return fut

    def close(self):
        with self._cv:
            self._closed = True
            self._cv.notify_all()
        for t in self._threads:
            t.join()
        for tpl in self._cache.values():
            if tpl.pid >= 0:
                _lib.call("mo_template_destroy", tpl.pid)
        self._cache.clear()

    def cache_state(self):
        """(resident template keys in LRU order, their bytes, budget)."""
        with self._cv:
            return list(self._cache.keys()), self._cache_bytes, self.budget

    # -- internals (the lock is held where noted) --------------------------
    def _fits(self, need):
        return self._cache_bytes + self._inflight_bytes + need <= self.budget

    def _idle(self, N=None):
        """Least recently used idle, built template (of grid N if given)."""
        return next((k for k, t in self._cache.items()
                     if t.users == 0 and t.ready.is_set() and t.error is None and (N is None or t.N == N)), None)

    def _evict_for(self, need):
        """Free idle templates, least recently used first, until `need` more
        bytes fit the budget (their ids are destroyed outside the lock).
        Lock held."""
        while not self._fits(need):
            victim = self._idle()
            if victim is None:
                return False
            tpl = self._cache.pop(victim)
            self._cache_bytes -= tpl.nbytes
            self.evictions += 1
            if tpl.pid >= 0:
                self._doomed.append(tpl.pid)
        return True

    def _reserve_template(self, key, N, need):
        """A template slot for a missing key: a new one when it fits the
        budget, else the least recently used idle one of the same grid
        resolution (rebuilt in place), else a new one after freeing idle
        templates.  Returns the _Template or None.  Lock held."""
        nbytes = template_bytes(N)
        if not self._fits(need + nbytes):
            victim = self._idle(N)
            if victim is not None:
                old = self._cache.pop(victim)
                self.evictions += 1
                tpl = _Template(N, old.nbytes, pid=old.pid)
                self._cache[key] = tpl
                return tpl
            if not self._evict_for(need + nbytes):
                return None
        tpl = _Template(N, nbytes)
        self._cache[key] = tpl
        self._cache_bytes += nbytes
        return tpl

    def _take_batch(self, cls):
        """Admit the longest FIFO prefix of the class's queue that fits the
        budget, pinning each request's template (reserved for a build when
        it is missing).  A request that fits nothing runs alone once no
        other work holds memory.  Lock held.  Returns
        [(req, fut, need, tpl, builder)]."""
        q = self._queues[cls]
        batch = []
        while q and len(batch) < self.max_batch:
            req, fut = q[0]
            need = solver_bytes(len(req.src_V), len(req.src_F))
            key = (req.tgt_key, req.grid)
            tpl = self._cache.get(key)
            builder = False
            if tpl is None:
                tpl = self._reserve_template(key, req.grid, need)
                if tpl is None:
                    if batch or self._inflight_bytes > 0:
                        break
                    tpl = _Template(req.grid, template_bytes(req.grid))   # oversized: alone
                    self._cache[key] = tpl
                    self._cache_bytes += tpl.nbytes
                builder = True
            elif not self._evict_for(need) and (batch or self._inflight_bytes > 0):
                break
            q.popleft()
            self._cache.move_to_end(key)
            tpl.users += 1
            self._inflight_bytes += need
            batch.append((req, fut, need, tpl, builder))
        return batch

    def _build(self, req, tpl):
        """Build a reserved template, or rebuild a recycled one in place (no
        lock held)."""
        from . import pyDeform
        try:
            if tpl.pid >= 0:
                deform.template_rebuild_f64(tpl.pid, req.tgt_V, req.tgt_F)
            else:
                tpl.pid = deform.template_f64(req.tgt_V, req.tgt_F, req.grid)
            _, tpl.scale, tpl.pos = pyDeform.TemplateInfo(tpl.pid)
        except BaseException as e:  # noqa: BLE001 - waiters get the same error
            tpl.error = e
        finally:
            tpl.ready.set()

    def _worker(self, cls):
        stream = torch.cuda.Stream()
        while True:
            with self._cv:
                while True:
                    while not self._queues[cls] and not self._closed:
                        self._cv.wait()
                    if self._closed and not self._queues[cls]:
                        return
                    batch = self._take_batch(cls)
                    if batch:
                        break
                    self._cv.wait()          # memory frees up when another batch ends
                doomed, self._doomed = self._doomed, []
            for pid in doomed:
                _lib.call("mo_template_destroy", pid)
            built = False
            try:
                with torch.cuda.stream(stream):
                    for req, _, _, tpl, builder in batch:
                        if builder:
                            recycled = tpl.pid >= 0
                            self._build(req, tpl)
                            built = True
                            with self._cv:
                                self.builds += 1
                                self.rebuilds += int(recycled)
                    good = []
                    for item in batch:
                        item[3].ready.wait()
                        if item[3].error is not None:      # its target's build failed: only it fails
                            item[1].set_exception(item[3].error)
                        else:
                            good.append(item)
                    Vs, Fs = [], []
                    for req, _, _, tpl, _ in good:
                        x0 = (np.asarray(req.src_V, np.float64) - tpl.pos) / tpl.scale    # Mesh::ApplyTransform
                        Vs.append(torch.from_numpy(x0).to("cuda"))
                        Fs.append(np.asarray(req.src_F, np.int32))
                    stats = deform.deform_batch([b[3].pid for b in good], Vs, Fs, [0] * len(good),
                                                [b[0].lam for b in good], options=self.options)
                    outs = [V.cpu().numpy() for V in Vs]
                done = time.perf_counter()
                with self._cv:
                    self.batches.append((cls, len(good)))
                    self.completed += len(good)
                for (req, fut, _, tpl, _), x, st in zip(good, outs, stats):
                    fut.set_result(DeformResult(x * tpl.scale + tpl.pos, st, done - req.arrival, len(good), built))
            except BaseException as e:  # noqa: BLE001 - surface every failure to the callers
                for _, fut, _, _, _ in batch:
                    if not fut.done():
                        fut.set_exception(e)
            finally:
                with self._cv:
                    for req, _, need, tpl, _ in batch:
                        tpl.users -= 1
                        self._inflight_bytes -= need
                        if tpl.error is not None:
                            key = (req.tgt_key, req.grid)
                            if self._cache.get(key) is tpl:
                                self._cache.pop(key)
                                self._cache_bytes -= tpl.nbytes
                                if tpl.pid >= 0:          # a recycled slot whose rebuild failed
                                    self._doomed.append(tpl.pid)
                    self._cv.notify_all()
