"""Threshold estimation for Rips persistence (reference threshold.py).

- ``auto_threshold``: percentile of k-nearest-neighbour distances.
- ``enclosing_radius``: min over points of the max distance to any other
  point; simplices with a larger diameter cannot carry a non-trivial pair.

Both run on the GPU through libflashph (fph_knn_threshold_batch,
fph_enclosing_radius_batch) -- CPU tensors on the host pipeline
(libflashph_cpu, same arithmetic): one thread per query point scans its cloud
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
    k = max(1, min(k, n - 1))
    if subsample is not None and subsample >= 2 and subsample < n:
        idx = torch.randperm(n, device=pts.device)[:subsample]
        pts = pts[idx]
        n = pts.shape[0]
        k = max(1, min(k, n - 1))
    percentile = max(0.0, min(100.0, percentile))
    if pts.device.type != "cuda":   # host memory: the host pipeline
        from flash_ph import _cpu
        kth = _cpu.knn_kth_dsq(pts.contiguous().numpy(), k)
        kth = np.sqrt(kth.astype(np.float64)).astype(np.float32)  # correctly rounded
        return float(_quantile_f32(kth, percentile / 100.0))
    _lib.require_gpu()
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
    if subsample is not None and subsample >= 2 and subsample < n:
        idx = torch.randperm(n, device=pts.device)[:subsample]
        pts = pts[idx]
    if pts.device.type != "cuda":   # host memory: the host pipeline
        from flash_ph import _cpu
        far = _cpu.far_dsq(pts.contiguous().numpy())
        return float(np.float32(np.sqrt(np.float64(far.min()))))
    _lib.require_gpu()
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
    """float32 tensor (threshold.py:123-130); NumPy input goes to the GPU
    when one is visible, as rips_persistence's."""
    if isinstance(points, np.ndarray):
        points = torch.from_numpy(points)
        if torch.cuda.is_available():
            points = points.cuda()
    return points.float()


def _quantile_f32(values: np.ndarray, q: float) -> np.float32:
    """torch.quantile's linear interpolation in FP32 (threshold.py:77):
    rank = q (n - 1) in FP32, then torch's lerp between the order statistics
    below and above (a + w (b - a) for w < 0.5, else b - (b - a)(1 - w)) --
    the formula of the GPU kernel k_quantile."""
    v = np.sort(np.asarray(values, dtype=np.float32))
    rank = np.float32(np.float32(q) * np.float32(v.size - 1))
    lo, hi = int(np.floor(rank)), int(np.ceil(rank))
    w = np.float32(rank - np.float32(lo))
    a, b = v[lo], v[hi]
    if w < np.float32(0.5):
        return np.float32(a + np.float32(w * np.float32(b - a)))
    return np.float32(b - np.float32(np.float32(b - a) * np.float32(np.float32(1.0) - w)))
