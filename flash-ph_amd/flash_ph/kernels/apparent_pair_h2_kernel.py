"""Ripser-style apparent-pair scan for H2 (fph_apparent_scan_h2).

Same call as the reference's h2_apparent_pair_scan
(kernels/apparent_pair_h2_kernel.py:328-441): for each triangle
(a, b, c), a < b < c, of largest edge rank r_t, the youngest tetrahedron
cofacet over the common neighbours of a, b and c; the triangle is apparent
when that cofacet's largest edge rank is r_t and the triangle is its oldest
face (the added vertex d below the vertex s off the largest edge).  The
pipeline's H2 stage uses the clearing-aware predicate of fph_apparent on
the uncleared triangles; on those, this set is a subset of it.
"""
from __future__ import annotations

import torch
from torch import Tensor

from flash_ph import _lib


def h2_apparent_pair_scan(
    tri_v0: Tensor,
    tri_v1: Tensor,
    tri_v2: Tensor,
    tri_max_rank: Tensor,
    adj_ptr: Tensor,
    adj_idx: Tensor,
    adj_rank: Tensor,
    max_degree: int,
    *,
    complete: bool = False,
) -> tuple[Tensor, Tensor]:
    """Classify triangles as apparent pairs for H2 (reference
    kernels/apparent_pair_h2_kernel.py:328-384).

    Returns (is_apparent (T,) int32, pivot_colex (T,) int64: the
    tetrahedron's colex index, 0 where not apparent) on the inputs' device.
    CUDA tensors: ``max_degree`` bounds the merge walk as the reference's GPU
    loop does (3 md steps, md = max_degree rounded up to a power of two in
    [64, 2048]); ``complete=True`` lifts the bound.  CPU tensors: the
    complete scan on the host (the reference's Numba path ignores max_degree
    too).
    """
    T = tri_v0.shape[0]
    device = tri_v0.device
    if T == 0:
        return (torch.zeros(0, dtype=torch.int32, device=device),
                torch.zeros(0, dtype=torch.int64, device=device))
    if device.type != "cuda":
        from flash_ph import _cpu

        def host(t):
            return t.detach().to(torch.int32).contiguous().numpy()

        flag, colex = _cpu.apparent_scan_h2(host(tri_v0), host(tri_v1), host(tri_v2),
                                            host(tri_max_rank), host(adj_ptr), host(adj_idx),
                                            host(adj_rank))
        return torch.from_numpy(flag), torch.from_numpy(colex)
    _lib.require_gpu()

    def prep(t):
        return t.to(device=device, dtype=torch.int32).contiguous()

    a, b, c, r = prep(tri_v0), prep(tri_v1), prep(tri_v2), prep(tri_max_rank)
    ap, ai, ar = prep(adj_ptr), prep(adj_idx), prep(adj_rank)
    flag = torch.empty(T, dtype=torch.int32, device=device)
    colex = torch.empty(T, dtype=torch.int64, device=device)
    lib = _lib.load()
    md = 0 if complete else max(int(max_degree), 1)  # as in apparent_pair_scan
    _lib.check(lib.fph_apparent_scan_h2(a.data_ptr(), b.data_ptr(), c.data_ptr(), r.data_ptr(),
                                        T, ap.data_ptr(), ai.data_ptr(), ar.data_ptr(),
                                        md, flag.data_ptr(), colex.data_ptr(),
                                        _lib.stream_of(device)), "fph_apparent_scan_h2")
    return flag, colex
