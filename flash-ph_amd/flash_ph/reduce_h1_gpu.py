"""GPU-native H1 reduction (reference reduce_h1_gpu.py).

Columns are edges, cofacets are triangles (cohomology, ripser convention).
Triangles come out of the enumeration already in filtration order, apparent
pairs are detected per column, and the residual columns are reduced on the
GPU (libflashph stages c, d, e).
"""
from __future__ import annotations

import torch

from flash_ph import _lib, _pipeline

_VALID_BACKENDS = ('auto', 'gpu', 'giotto')


def rips_h1_gpu_native(filt, n, mst_idx, max_edge_length, device, backend='auto'):
    """H1 diagram of a filtration (reference reduce_h1_gpu.py:438-604).

    ``filt`` is a :class:`flash_ph.rips.RipsFiltration` on the GPU and
    ``mst_idx`` the forest edge indices (``gpu_boruvka_mst``).  Returns a
    (K, 2) float32 tensor in the reference's row order.
    """
    if backend not in _VALID_BACKENDS:
        raise ValueError(
            f"backend must be one of {_VALID_BACKENDS}, got {backend!r}"
        )
    if backend == 'giotto':
        from flash_ph.persistence import _rips_h1h2_giotto
        return _rips_h1h2_giotto(filt, n, 1, max_edge_length, device)[0]
    _lib.require_gpu()
    E = filt.edge_i.shape[0]
    cx = _pipeline.Complex(n=n, device=filt.edge_i.device)
    cx.E = E
    cx.edge_i = filt.edge_i.to(torch.int32).contiguous()
    cx.edge_j = filt.edge_j.to(torch.int32).contiguous()
    cx.edge_dsq = filt.edge_dist_sq.contiguous()
    cx.edge_filt = torch.sqrt(filt.edge_dist_sq).contiguous()
    cx.adj_ptr = filt.vert_adj_ptr.contiguous()
    cx.adj_idx = filt.vert_adj_idx.contiguous()
    cx.adj_rank = filt.vert_adj_rank.contiguous()
    cx.adj_dsq = filt.vert_adj_dist_sq.contiguous()
    mask = torch.zeros(max(E, 1), dtype=torch.uint8, device=cx.device)
    if mst_idx.numel() > 0:
        mask[mst_idx.long()] = 1
    cx.mst_mask = mask
    cx.info["converged"] = torch.ones(1, dtype=torch.int32, device=cx.device)
    # 'auto' and 'gpu': the reference's store semantics (see rips_persistence)
    _pipeline.build_h1(cx, semantics="reference")
    return cx.h1.to(device)
