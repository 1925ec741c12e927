"""Batches of point clouds of uneven size and dimension in one pass.

The reference handles many clouds by calling ``rips_persistence`` (and
``auto_threshold``) once per cloud (persistence.py:68-206,
threshold.py:20-78).  Here the clouds are concatenated into one device
buffer and every stage runs once over all of them: per-cloud screening
against per-cloud thresholds, one filtration sort grouped by cloud, one
forest, one triangle/cofacet pass and one reduction with a workgroup per
cloud.  Every cloud's diagrams are bit-identical to what
``rips_persistence`` returns for that cloud alone with the same threshold.
"""
from __future__ import annotations

from collections.abc import Sequence

import numpy as np
import torch
from torch import Tensor

from flash_ph import _lib, _pipeline


def _prepare(clouds) -> tuple[list, torch.device, torch.device]:
    if isinstance(clouds, (np.ndarray, Tensor)) and clouds.ndim == 3:
        clouds = list(clouds)
    if not isinstance(clouds, Sequence) or isinstance(clouds, (str, bytes)):
        raise TypeError("clouds must be a sequence of (n, d) point arrays")
    out_dev = None
    prepared = []
    for i, c in enumerate(clouds):
        from_numpy = isinstance(c, np.ndarray)
        if from_numpy:
            c = torch.from_numpy(c)
        if not isinstance(c, Tensor):
            raise TypeError(f"cloud {i}: expected a numpy array or tensor, got {type(c)}")
        if c.ndim != 2:
            raise ValueError(f"cloud {i}: points must have shape (n, d), got {tuple(c.shape)}")
        if c.shape[1] < 1 and c.shape[0] > 0:
            raise ValueError(f"cloud {i}: points must have d >= 1")
        if c.shape[0] > _pipeline.MAX_BATCH_CLOUD:
            raise ValueError(
                f"cloud {i}: n={c.shape[0]} exceeds {_pipeline.MAX_BATCH_CLOUD}, the largest "
                f"cloud a batch keeps in the reference's edge order; use rips_persistence")
        if out_dev is None and not from_numpy:
            out_dev = c.device
        prepared.append(c)
    _lib.require_gpu()
    dev = torch.device("cuda", torch.cuda.current_device())
    for p in prepared:
        if p.device.type == "cuda":
            dev = p.device
            break
    prepared = [p.to(dev).float().contiguous() for p in prepared]
    return prepared, dev, (out_dev if out_dev is not None else dev)


def auto_threshold_batch(clouds, k: int = 20, percentile: float = 95.0) -> list[float]:
    """``auto_threshold`` of every cloud in one launch.

    Each value is the ``percentile`` of the clouds' k-th nearest-neighbour
    distances (k clamped to [1, n-1]; 0.0 for n <= 1), with ``torch.quantile``'s
    linear interpolation in FP32.  Distances are the filtration's own FP32
    values (the edge screen's FMA chain, correctly rounded sqrt); the
    reference's cdist-based values agree to within cdist's rounding.
    """
    pts, dev, _ = _prepare(clouds)
    if not pts:
        return []
    L = _pipeline.layout_batch(pts, dev)
    t, _ = _pipeline.knn_threshold_batch(L, max(1, int(k)), float(percentile))
    return [float(x) for x in t.tolist()]


def rips_persistence_batch(
    clouds,
    max_edge_length=None,
    max_dim: int = 1,
    k: int = 20,
    percentile: float = 95.0,
    backend: str = "auto",
) -> list[list[Tensor]]:
    """``rips_persistence`` of many clouds in one pass.

    Parameters
    ----------
    clouds : sequence of (n_c, d_c) arrays or tensors
        Clouds of any sizes and dimensions (n_c <= 92682).
    max_edge_length : float, sequence of floats, or None
        One threshold for all clouds, one per cloud, or None for the
        per-cloud ``auto_threshold(cloud, k, percentile)`` computed on the
        device (see ``auto_threshold_batch``).  Strict < as the reference.
    max_dim : int
        0 (H0) or 1 (H0 + H1).
    backend : str
        ``'auto'`` / ``'gpu'``: the reference's reduction semantics (as
        ``rips_persistence``); ``'exact'``: the parallel complete reduction.

    Returns
    -------
    list of [H0, H1] (or [H0]) per cloud, each (K, 2) float32, in the order
    and with the values ``rips_persistence(cloud, t_c, max_dim)`` gives.
    """
    if max_dim not in (0, 1):
        raise ValueError(f"batched max_dim must be 0 or 1, got {max_dim}")
    if backend not in ("auto", "gpu", "exact"):
        raise ValueError(f"backend must be one of ('auto', 'gpu', 'exact'), got {backend!r}")
    pts, dev, out_dev = _prepare(clouds)
    C = len(pts)
    if C == 0:
        return []
    L = _pipeline.layout_batch(pts, dev)
    if max_edge_length is None:
        _, thr_sq = _pipeline.knn_threshold_batch(L, max(1, int(k)), float(percentile))
    else:
        if np.ndim(max_edge_length) == 0:
            ts = [float(max_edge_length)] * C
        else:
            ts = [float(x) for x in max_edge_length]
            if len(ts) != C:
                raise ValueError(f"got {len(ts)} thresholds for {C} clouds")
        for t in ts:
            if t < 0:
                raise ValueError(f"max_edge_length must be non-negative, got {t}")
        thr_sq = torch.tensor([_pipeline.threshold_sq_f32(t) for t in ts], dtype=torch.float32,
                              device=dev)
    semantics = "exact" if backend == "exact" else "reference"
    if torch.is_grad_enabled() and any(getattr(c, "requires_grad", False) for c in clouds):
        from flash_ph.autograd import rips_persistence_batch_diff
        dgms = rips_persistence_batch_diff(pts, thr_sq, max_dim, semantics)
    else:
        dgms, _ = _pipeline.compute_batch(L, thr_sq, max_dim, semantics=semantics)
    if out_dev != dev:
        dgms = [[d.to(out_dev) for d in per] for per in dgms]
    return dgms
