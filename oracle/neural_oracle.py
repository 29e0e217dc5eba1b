"""CPU restatement of the NeuralODE bidirectional driver's native pieces.

TEST INFRASTRUCTURE ONLY: imported by tests/ as the checker.  The product
path (meshode_amd) never imports this module.

Restates:
  * ReverseLossLayer's nearest-neighbour query (reverse_loss_layer.py:15-19,
    scipy cKDTree(src).query(tar, k=1)) as a brute force: squared distances
    in FP64 from the FP32 coordinates, (dx*dx + dy*dy) + dz*dz, lowest
    index among equal distances; and its loss 1/2 sum |src[nn] - tar|^2
    (:21-24) in FP64.
"""
from __future__ import annotations

import numpy as np


def nearest(src, tar, block=2048):
    """(index int64 [nt], squared distance float64 [nt])."""
    S = np.asarray(src, np.float32).astype(np.float64)
    T = np.asarray(tar, np.float32).astype(np.float64)
    idx = np.empty(len(T), np.int64)
    d2 = np.empty(len(T))
    for lo in range(0, len(T), block):
        t = T[lo:lo + block]
        dx = S[None, :, 0] - t[:, None, 0]
        dy = S[None, :, 1] - t[:, None, 1]
        dz = S[None, :, 2] - t[:, None, 2]
        D = (dx * dx + dy * dy) + dz * dz
        k = np.argmin(D, axis=1)              # first occurrence: lowest index
        idx[lo:lo + block] = k
        d2[lo:lo + block] = D[np.arange(len(t)), k]
    return idx, d2


def reverse_loss(src, tar):
    """1/2 sum |src[nn] - tar|^2 in FP64."""
    S = np.asarray(src, np.float64)
    T = np.asarray(tar, np.float64)
    k, _ = nearest(src, tar)
    d = S[k] - T
    return 0.5 * float((d * d).sum())
