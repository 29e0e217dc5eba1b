"""Continuous-batching deformation server on one GPU.

Requests (source mesh + target mesh or a cached target key, grid resolution,
rigidity) arrive at any time.  Worker threads, each with its own HIP stream,
repeatedly take every waiting request that fits the memory budget, make sure
each target's template (normalization + distance grid) is resident, and run
the whole group through one batched Levenberg-Marquardt launch
(Deformer::Deform, deformer.cc:18-92).  While one worker's batch is on the
GPU the next worker gathers and launches the requests that arrived
meanwhile, so batches overlap on the device.

Target templates are kept in HBM in an LRU cache keyed by the request's
target key; admission charges each request its solver working set and each
newly built template its grid bytes, and evicts least recently used idle
templates when the budget is exceeded.  A request's result does not depend on
which other requests share its batch (each pair is solved by its own
workgroup with a fixed reduction order).
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


def solver_bytes(nv, nf):
    """Device working set of one pair in the batched solver (bytes)."""
    return nv * (2 * 16 + 8 + 4 + 13 * 16 + 24) + 3 * nf * 2 * 4 * 2


def template_bytes(N, nv_t, nf_t):
    return N ** 3 * (8 + 4 + 4) + nv_t * 24 + nf_t * 12


class _Template:
    __slots__ = ("pid", "scale", "pos", "nbytes", "users")

    def __init__(self, pid, scale, pos, nbytes):
        self.pid, self.scale, self.pos, self.nbytes, self.users = pid, scale, pos, nbytes, 0


class DeformServer:
    def __init__(self, mem_budget_bytes=200e9, max_batch=512, workers=2, options=None):
        if not torch.cuda.is_available():
            raise RuntimeError("DeformServer needs a ROCm GPU")
        _lib.load()
        self.budget = float(mem_budget_bytes)
        self.max_batch = int(max_batch)
        self.options = options or deform.default_options()
        self._cv = threading.Condition()
        self._queue = collections.deque()
        self._cache = collections.OrderedDict()       # key -> _Template (LRU order)
        self._cache_bytes = 0.0
        self._inflight_bytes = 0.0
        self._closed = False
        self.completed = 0
        self.batches = []
        self._threads = [threading.Thread(target=self._worker, args=(i,), daemon=True) for i in range(workers)]
        for t in self._threads:
            t.start()

    # -- public ----------------------------------------------------------
    def submit(self, req: DeformRequest) -> Future:
        fut = Future()
        with self._cv:
            if self._closed:
                raise RuntimeError("server closed")
            self._queue.append((req, fut))
            self._cv.notify()
        return fut

    def close(self):
        with self._cv:
            self._closed = True
            self._cv.notify_all()
        for t in self._threads:
            t.join()
        for tpl in self._cache.values():
            _lib.call("mo_template_destroy", tpl.pid)
        self._cache.clear()

    # -- internals -------------------------------------------------------
    def _take_batch(self):
        """Pop the longest FIFO prefix of waiting requests that fits the budget
        (at least one request, so an oversized request still runs alone)."""
        batch = []
        while self._queue and len(batch) < self.max_batch:
            req, fut = self._queue[0]
            need = solver_bytes(len(req.src_V), len(req.src_F))
            if (req.tgt_key, req.grid) not in self._cache:
                need += template_bytes(req.grid, len(req.tgt_V), len(req.tgt_F))
            if batch and self._inflight_bytes + self._cache_bytes + need > self.budget:
                break
            self._queue.popleft()
            self._inflight_bytes += need
            batch.append((req, fut, need))
        return batch

    def _evict(self):
        while self._cache_bytes + self._inflight_bytes > self.budget:
            victim = next((k for k, t in self._cache.items() if t.users == 0), None)
            if victim is None:
                return
            tpl = self._cache.pop(victim)
            self._cache_bytes -= tpl.nbytes
            _lib.call("mo_template_destroy", tpl.pid)

    def _template(self, req):
        """Cached template for the request's target (built on a miss); the
        caller holds the lock."""
        key = (req.tgt_key, req.grid)
        tpl = self._cache.get(key)
        if tpl is None:
            pid = deform.template_f64(req.tgt_V, req.tgt_F, req.grid)
            N, scale, pos = self._info(pid)
            tpl = _Template(pid, scale, pos, template_bytes(req.grid, len(req.tgt_V), len(req.tgt_F)))
            self._cache[key] = tpl
            self._cache_bytes += tpl.nbytes
        self._cache.move_to_end(key)
        tpl.users += 1
        return tpl

    @staticmethod
    def _info(pid):
        from . import pyDeform
        return pyDeform.TemplateInfo(pid)

    def _worker(self, wid):
        stream = torch.cuda.Stream()
        while True:
            with self._cv:
                while not self._queue and not self._closed:
                    self._cv.wait()
                if self._closed and not self._queue:
                    return
                batch = self._take_batch()
            tpls = []
            try:
                with torch.cuda.stream(stream):
                    with self._cv:
                        for req, _, _ in batch:
                            tpls.append(self._template(req))
                        self._evict()
                    Vs, Fs = [], []
                    for (req, _, _), tpl in zip(batch, tpls):
                        x0 = (np.asarray(req.src_V, np.float64) - tpl.pos) / tpl.scale    # Mesh::ApplyTransform
                        Vs.append(torch.from_numpy(x0).to("cuda"))
                        Fs.append(np.asarray(req.src_F, np.int32))
                    stats = deform.deform_batch([t.pid for t in tpls], Vs, Fs, [0] * len(batch),
                                                [req.lam for req, _, _ in batch], options=self.options)
                    outs = [V.cpu().numpy() for V in Vs]
                done = time.perf_counter()
                with self._cv:
                    self.batches.append(len(batch))
                    self.completed += len(batch)
                for (req, fut, _), tpl, x, st in zip(batch, tpls, outs, stats):
                    fut.set_result(DeformResult(x * tpl.scale + tpl.pos, st, done - req.arrival, len(batch)))
            except BaseException as e:  # noqa: BLE001 - surface every failure to the callers
                for _, fut, _ in batch:
                    if not fut.done():
                        fut.set_exception(e)
            finally:
                with self._cv:
                    for t in tpls:
                        t.users -= 1
                    for _, _, need in batch:
                        self._inflight_bytes -= need
