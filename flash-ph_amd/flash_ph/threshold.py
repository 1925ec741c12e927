"""Threshold estimation for Rips persistence (reference threshold.py).

- ``auto_threshold``: percentile of k-nearest-neighbour distances.
- ``enclosing_radius``: min over points of the max distance to any other
  point; simplices with a larger diameter cannot carry a non-trivial pair.

Both follow the reference's tiled torch formulation (threshold.py:133-200):
doubly tiled ``torch.cdist`` with a running top-k, then ``torch.quantile``.
"""
from __future__ import annotations

import numpy as np
import torch
from torch import Tensor


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
    k = max(1, min(k, n - 1))
    if subsample is not None and subsample >= 2 and subsample < n:
        idx = torch.randperm(n, device=pts.device)[:subsample]
        pts = pts[idx]
        n = pts.shape[0]
        k = max(1, min(k, n - 1))
    percentile = max(0.0, min(100.0, percentile))
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
    if subsample is not None and subsample >= 2 and subsample < n:
        idx = torch.randperm(n, device=pts.device)[:subsample]
        pts = pts[idx]
    return float(_tiled_enclosing_radius(pts))


def _to_tensor(points) -> Tensor:
    if isinstance(points, np.ndarray):
        t = torch.from_numpy(points).float()
        if torch.cuda.is_available():
            t = t.cuda()
        return t
    return points.float()


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
