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
For k > 128 auto_threshold keeps the reference's tiled cdist + top-k
formulation, on the GPU.
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
    if k <= _pipeline.KNN_MAX_K:
        L = _pipeline.layout_batch([pts.contiguous()], pts.device)
        t, _ = _pipeline.knn_threshold_batch(L, k, percentile)
        return float(t.item())
    knn_dists = _tiled_knn(pts, k)
    return float(torch.quantile(knn_dists, percentile / 100.0).item())


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


def _tiled_knn(pts: Tensor, k: int, tile: int = 4096) -> Tensor:
    """k-th nearest-neighbour distance per point via tiled cdist and a
    running top-k (reference threshold.py:133-176)."""
    n = pts.shape[0]
    device = pts.device
    topk_buf = torch.full((n, k), float('inf'), device=device)
    for i_start in range(0, n, tile):
        i_end = min(i_start + tile, n)
        qi = pts[i_start:i_end]
        for j_start in range(0, n, tile):
            j_end = min(j_start + tile, n)
            dist = torch.cdist(qi, pts[j_start:j_end])
            lo, hi = max(i_start, j_start), min(i_end, j_end)
            if lo < hi:
                r = torch.arange(lo, hi, device=device)
                dist[r - i_start, r - j_start] = float('inf')
            k_tile = min(k, j_end - j_start)
            tile_topk, _ = dist.topk(k_tile, dim=1, largest=False)
            combined = torch.cat([topk_buf[i_start:i_end], tile_topk], dim=1)
            topk_buf[i_start:i_end], _ = combined.topk(k, dim=1, largest=False)
    return topk_buf[:, -1]


def _tiled_enclosing_radius(pts: Tensor, tile: int = 4096) -> float:
    """min_x max_y d(x, y) via doubly tiled cdist (reference threshold.py:179-200)."""
    n = pts.shape[0]
    max_dists = torch.zeros(n, device=pts.device)
    for j_start in range(0, n, tile):
        chunk = pts[j_start:min(j_start + tile, n)]
        for i_start in range(0, n, tile):
            i_end = min(i_start + tile, n)
            tile_max, _ = torch.cdist(pts[i_start:i_end], chunk).max(dim=1)
            max_dists[i_start:i_end] = torch.maximum(max_dists[i_start:i_end], tile_max)
    return float(max_dists.min().item())
