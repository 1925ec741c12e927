"""Vietoris-Rips filtration construction (reference rips.py).

``rips_filtration`` screens all pairs on the GPU (libflashph stage a), sorts
the edges by the reference's packed key and builds the per-vertex CSR.
"""
from __future__ import annotations

from typing import NamedTuple

import torch
from torch import Tensor

from flash_ph import _lib, _pipeline


class RipsFiltration(NamedTuple):
    """Result of Rips filtration construction (reference rips.py:15-27).

    All edge arrays are sorted by filtration order (distance², then colex).
    """
    edge_i: Tensor        # (E,) int32 — source vertices (i < j)
    edge_j: Tensor        # (E,) int32 — target vertices (i < j)
    edge_dist_sq: Tensor  # (E,) float32 — squared edge lengths
    vert_adj_ptr: Tensor  # (n+1,) int32 — CSR row pointers
    vert_adj_idx: Tensor  # (2E,) int32 — neighbor vertex indices
    vert_adj_rank: Tensor # (2E,) int32 — edge_rank per adjacency entry
    vert_adj_dist_sq: Tensor  # (2E,) float32 — dist² per adjacency entry
    n: int                # number of original points


def _pack_sort_key(dist_sq: Tensor, edge_i: Tensor, edge_j: Tensor) -> Tensor:
    """(float_bits(dist²) << 32) | (colex mod 2**32), colex = j(j-1)/2 + i.

    Same key as the reference (rips.py:30-50); the GPU path builds it inside
    the screening kernel.  Kept for callers that pack keys themselves.
    """
    dist_bits = dist_sq.view(torch.int32).to(torch.int64) & 0xFFFFFFFF
    j_long = edge_j.long()
    i_long = edge_i.long()
    colex = j_long * (j_long - 1) // 2 + i_long
    return (dist_bits << 32) | (colex & 0xFFFFFFFF)


def _empty_filtration(n: int, device) -> RipsFiltration:
    e32 = torch.empty(0, dtype=torch.int32, device=device)
    ef = torch.empty(0, dtype=torch.float32, device=device)
    return RipsFiltration(edge_i=e32, edge_j=e32.clone(), edge_dist_sq=ef,
                          vert_adj_ptr=torch.zeros(n + 1, dtype=torch.int32, device=device),
                          vert_adj_idx=e32.clone(), vert_adj_rank=e32.clone(),
                          vert_adj_dist_sq=ef.clone(), n=n)


def _build_vertex_csr(edge_i: Tensor, edge_j: Tensor, edge_rank: Tensor,
                      edge_dist_sq: Tensor, n: int):
    """Per-vertex CSR with (neighbor, edge_rank, dist²) (reference
    rips.py:53-95).  Edge ranks are the positions in the given (sorted) edge
    order; a non-identity ``edge_rank`` is applied by reordering the edges."""
    _lib.require_gpu()
    device = edge_i.device
    E = edge_i.shape[0]
    if E == 0:
        f = _empty_filtration(n, device)
        return f.vert_adj_ptr, f.vert_adj_idx, f.vert_adj_rank, f.vert_adj_dist_sq
    order = torch.argsort(edge_rank.long())
    ei = edge_i[order].to(torch.int32).contiguous()
    ej = edge_j[order].to(torch.int32).contiguous()
    ed = edge_dist_sq[order].to(torch.float32).contiguous()
    lo, hi = torch.minimum(ei, ej), torch.maximum(ei, ej)
    ptr_ = torch.empty(n + 1, dtype=torch.int32, device=device)
    idx = torch.empty(2 * E, dtype=torch.int32, device=device)
    rank = torch.empty(2 * E, dtype=torch.int32, device=device)
    dsq = torch.empty(2 * E, dtype=torch.float32, device=device)
    _lib.call_ws("fph_vertex_csr",
                 (_lib.ptr(lo.contiguous()), _lib.ptr(hi.contiguous()), _lib.ptr(ed), E, n,
                  _lib.ptr(ptr_), _lib.ptr(idx), _lib.ptr(rank), _lib.ptr(dsq)),
                 (_lib.stream_of(device),), device)
    # positions refer to the sorted order; map back to the caller's ranks
    rank = edge_rank.to(torch.int32)[order][rank.long()]
    return ptr_, idx, rank, dsq


def rips_filtration(points: Tensor, max_edge_length: float) -> RipsFiltration:
    """Construct a Vietoris-Rips filtration (reference rips.py:98-218): on the
    GPU for CUDA tensors, on the host pipeline for CPU tensors.

    Parameters
    ----------
    points : (n, d) Tensor
        Point cloud coordinates (any d >= 1).
    max_edge_length : float
        Maximum edge length (strict < inequality).

    Returns
    -------
    RipsFiltration
        Sorted edges and per-vertex CSR adjacency, on ``points.device``.
    """
    if max_edge_length < 0:
        raise ValueError(
            f"max_edge_length must be non-negative, got {max_edge_length}"
        )
    n, d = points.shape
    device = points.device
    if n <= 1:
        return _empty_filtration(n, device)
    if device.type != "cuda":
        # host memory: the host pipeline (the reference's CPU route,
        # rips.py:221-233, with the GPU's FMA-form distances)
        from flash_ph import _cpu
        h = _cpu.filtration(points.detach().float().contiguous().numpy(),
                            _pipeline.threshold_sq_f32(max_edge_length))
        if h["edge_i"].size == 0:
            return _empty_filtration(n, device)
        t = {k: torch.from_numpy(v) for k, v in h.items()}
        return RipsFiltration(edge_i=t["edge_i"], edge_j=t["edge_j"],
                              edge_dist_sq=t["edge_dsq"], vert_adj_ptr=t["adj_ptr"],
                              vert_adj_idx=t["adj_idx"], vert_adj_rank=t["adj_rank"],
                              vert_adj_dist_sq=t["adj_dsq"], n=n)
    _lib.require_gpu()
    pts = points.float().contiguous()
    if pts.device.type != "cuda":
        pts = pts.cuda()
    cx = _pipeline.build_edges(pts, max_edge_length)
    if cx.E == 0:
        return _empty_filtration(n, device)
    f = RipsFiltration(edge_i=cx.edge_i, edge_j=cx.edge_j, edge_dist_sq=cx.edge_dsq,
                       vert_adj_ptr=cx.adj_ptr, vert_adj_idx=cx.adj_idx,
                       vert_adj_rank=cx.adj_rank, vert_adj_dist_sq=cx.adj_dsq, n=n)
    if device.type != "cuda":
        f = RipsFiltration(*[t.to(device) if isinstance(t, Tensor) else t for t in f])
    return f
