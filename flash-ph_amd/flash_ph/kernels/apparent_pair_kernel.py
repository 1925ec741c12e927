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
    *,
    complete: bool = False,
) -> tuple[Tensor, Tensor]:
    """Classify non-MST edges as apparent (reference
    kernels/apparent_pair_kernel.py:200-236).

    Returns (is_apparent (K,) int32, pivot_colex (K,) int64) on the inputs'
    device.  CUDA tensors: ``max_degree`` bounds the merge walk as the
    reference's GPU loop does (md = max_degree rounded up to a power of two
    in [64, 2048], 2 md steps); ``complete=True`` lifts the bound.  CPU
    tensors: the complete scan on the host (the reference's Numba path
    ignores max_degree too).
    """
    K = nonmst_ranks.shape[0]
    device = edge_i.device
    if K == 0:
        return (torch.zeros(0, dtype=torch.int32, device=device),
                torch.zeros(0, dtype=torch.int64, device=device))
    if device.type != "cuda":
        from flash_ph import _cpu

        def host(t):
            return t.detach().to(torch.int32).contiguous().numpy()

        flag, colex = _cpu.apparent_scan(host(edge_i), host(edge_j), host(nonmst_ranks),
                                         host(adj_ptr), host(adj_idx), host(adj_rank))
        return torch.from_numpy(flag), torch.from_numpy(colex)
    _lib.require_gpu()

    def prep(t):
        return t.to(device=device, dtype=torch.int32).contiguous()

    ei, ej, nm = prep(edge_i), prep(edge_j), prep(nonmst_ranks)
    ap, ai, ar = prep(adj_ptr), prep(adj_idx), prep(adj_rank)
    flag = torch.empty(K, dtype=torch.int32, device=device)
    colex = torch.empty(K, dtype=torch.int64, device=device)
    lib = _lib.load()
    # the C entry point takes max_degree <= 0 as "no bound"; the reference's
    # loop treats every max_degree <= 64 as md = 64
    md = 0 if complete else max(int(max_degree), 1)
    _lib.check(lib.fph_apparent_scan(ei.data_ptr(), ej.data_ptr(), nm.data_ptr(), K,
                                     ap.data_ptr(), ai.data_ptr(), ar.data_ptr(), md,
                                     flag.data_ptr(), colex.data_ptr(), _lib.stream_of(device)),
               "fph_apparent_scan")
    return flag, colex
