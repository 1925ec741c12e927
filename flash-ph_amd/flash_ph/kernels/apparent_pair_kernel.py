"""Ripser-style apparent-pair scan for H1 (fph_apparent_scan).

Same call as the reference's apparent_pair_scan
(kernels/apparent_pair_kernel.py:200-313): for each non-forest edge
e = (u, v) of rank r_e, is there a common neighbour w with rank(u, w) < r_e
and rank(v, w) < r_e; if so, the colex index of the triangle with the
smallest such w.  The pipeline itself uses the forest-cleared predicate of
reduce_h1_gpu.py:159-223 (fph_apparent), a superset of this one.
"""
from __future__ import annotations

import torch
from torch import Tensor

from flash_ph import _lib


def apparent_pair_scan(
    edge_i: Tensor,
    edge_j: Tensor,
    nonmst_ranks: Tensor,
    adj_ptr: Tensor,
    adj_idx: Tensor,
    adj_rank: Tensor,
    max_degree: int,
) -> tuple[Tensor, Tensor]:
    """Classify non-MST edges as apparent (reference
    kernels/apparent_pair_kernel.py:200-236).

    Returns (is_apparent (K,) int32, pivot_colex (K,) int64) on the inputs'
    device.  ``max_degree`` bounds the merge walk as the reference's GPU
    loop does (md = max_degree rounded up to a power of two in [64, 2048],
    2 md steps); pass 0 for a complete scan (the reference's CPU path).
    CPU tensors are computed on the GPU and returned on the CPU.
    """
    K = nonmst_ranks.shape[0]
    device = edge_i.device
    if K == 0:
        return (torch.zeros(0, dtype=torch.int32, device=device),
                torch.zeros(0, dtype=torch.int64, device=device))
    _lib.require_gpu()
    dev = device if device.type == "cuda" else torch.device("cuda")

    def prep(t):
        return t.to(device=dev, dtype=torch.int32).contiguous()

    ei, ej, nm = prep(edge_i), prep(edge_j), prep(nonmst_ranks)
    ap, ai, ar = prep(adj_ptr), prep(adj_idx), prep(adj_rank)
    flag = torch.empty(K, dtype=torch.int32, device=dev)
    colex = torch.empty(K, dtype=torch.int64, device=dev)
    lib = _lib.load()
    _lib.check(lib.fph_apparent_scan(ei.data_ptr(), ej.data_ptr(), nm.data_ptr(), K,
                                     ap.data_ptr(), ai.data_ptr(), ar.data_ptr(),
                                     int(max_degree), flag.data_ptr(), colex.data_ptr(),
                                     _lib.stream_of(dev)), "fph_apparent_scan")
    if dev != device:
        return flag.to(device), colex.to(device)
    return flag, colex
