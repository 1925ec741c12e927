"""Threshold estimation for Rips persistence (reference threshold.py).

- ``auto_threshold``: percentile of k-nearest-neighbour distances.
- ``enclosing_radius``: min over points of the max distance to any other
  point; simplices with a larger diameter cannot carry a non-trivial pair.

Both run on the GPU through libflashph (fph_knn_threshold_batch,
fph_enclosing_radius_batch): one thread per query point scans its cloud
with the edge screen's FP32 FMA-chain distances, so a threshold derived
here is measured in the same arithmetic as the filtration it bounds.  The
reference computes the same quantities with torch.cdist (threshold.py:133-200),
whose GEMM-form rounding differs from the direct form in the last bits.
k <= 128 keeps each point's k smallest distances in registers; larger k
selects the k-th by radix over the distance bits.
"""
from __future__ import annotations

import numpy as np
import torch
from torch import Tensor

from flash_ph import _lib, _pipeline


def auto_threshold(
    points,
    k: int = 20,
    percentile: float = 95.0,
    subsample: int | None = None,
) -> float:
    """Estimate a Rips threshold from the k-NN distance distribution
    (reference threshold.py:20-78).

    Parameters
    ----------
    points : (n, d) array-like
    k : int
        Number of nearest neighbours.  Default 20.
    percentile : float
        Percentile of the k-NN distance distribution.  Default 95.
    subsample : int or None
        Randomly subsample this many points first.

    Returns
    -------
    float
        Estimated threshold (Euclidean distance).
    """
    pts = _to_tensor(points)
    n = pts.shape[0]
    if n <= 1:
        return 0.0
    pts = _on_gpu(pts)
    k = max(1, min(k, n - 1))
    if subsample is not None and subsample >= 2 and subsample < n:
        idx = torch.randperm(n, device=pts.device)[:subsample]
        pts = pts[idx]
        n = pts.shape[0]
        k = max(1, min(k, n - 1))
    percentile = max(0.0, min(100.0, percentile))
    L = _pipeline.layout_batch([pts.contiguous()], pts.device)
    t, _ = _pipeline.knn_threshold_batch(L, k, percentile)
    return float(t.item())


def enclosing_radius(
    points,
    subsample: int | None = None,
) -> float:
    """Enclosing radius min_x max_y d(x, y) (reference threshold.py:81-116)."""
    pts = _to_tensor(points)
    n = pts.shape[0]
    if n <= 1:
        return 0.0
    pts = _on_gpu(pts)
    if subsample is not None and subsample >= 2 and subsample < n:
        idx = torch.randperm(n, device=pts.device)[:subsample]
        pts = pts[idx]
    L = _pipeline.layout_batch([pts.contiguous()], pts.device)
    r = _enclosing_radius_batch(L)
    return float(r.item())


def _enclosing_radius_batch(L) -> Tensor:
    r = torch.empty(L.C, dtype=torch.float32, device=L.pts.device)
    _lib.call_ws("fph_enclosing_radius_batch",
                 (_lib.ptr(L.pts), _lib.ptr(L.foff), _lib.ptr(L.dims), _lib.ptr(L.vbase),
                  _lib.ptr(L.qtile_off), L.n_qtiles, L.C, max(1, L.max_d), _lib.ptr(r)),
                 (_lib.stream_of(L.pts.device),), L.pts.device)
    return r


def _to_tensor(points) -> Tensor:
    """float32 tensor (threshold.py:123-130)."""
    if isinstance(points, np.ndarray):
        points = torch.from_numpy(points)
    return points.float()


def _on_gpu(pts: Tensor) -> Tensor:
    """The computation runs on the GPU; CPU tensors are moved there."""
    _lib.require_gpu()
    return pts if pts.device.type == "cuda" else pts.cuda()
