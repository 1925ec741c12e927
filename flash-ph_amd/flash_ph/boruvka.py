"""GPU Boruvka minimum spanning forest for H0 (reference boruvka.py).

Edge weights are filtration ranks (distinct), so the forest is unique.
"""
from __future__ import annotations

import torch
from torch import Tensor

from flash_ph import _lib


def gpu_boruvka_mst(
    edge_i: Tensor,
    edge_j: Tensor,
    edge_rank: Tensor,
    n: int,
) -> tuple[Tensor, Tensor]:
    """Boruvka minimum spanning forest on the GPU (reference boruvka.py:17-113).

    Parameters
    ----------
    edge_i, edge_j : (E,) int32
        Edge endpoints, sorted by filtration.
    edge_rank : (E,) int32
        Edge ranks (weights); the reference passes ``arange(E)``.
    n : int
        Number of vertices.

    Returns
    -------
    mst_edge_indices : (K,) int64
        Indices into the edge list of the forest edges, ascending.
    final_comp : (n,) int64
        Component labels: the smallest vertex of each component.
    """
    _lib.require_gpu()
    lib = _lib.load()
    device = edge_i.device
    E = edge_i.shape[0]
    if n < 1:
        return (torch.empty(0, dtype=torch.int64, device=device),
                torch.empty(0, dtype=torch.int64, device=device))
    ident = E == 0 or bool(torch.equal(edge_rank.long().cpu(), torch.arange(E)))
    order = None if ident else torch.argsort(edge_rank.long())
    ei = (edge_i if order is None else edge_i[order]).to(torch.int32).contiguous()
    ej = (edge_j if order is None else edge_j[order]).to(torch.int32).contiguous()
    s = _lib.stream_of(device)
    mask = torch.empty(max(E, 1), dtype=torch.uint8, device=device)
    comp = torch.empty(n, dtype=torch.int64, device=device)
    conv = torch.empty(1, dtype=torch.int32, device=device)
    scratch = _lib.WORKSPACE.get(lib.fph_boruvka_scratch_bytes(n), device)
    _lib.check(lib.fph_boruvka_mst(_lib.ptr(ei), _lib.ptr(ej), E, n, _lib.ptr(mask),
                                   _lib.ptr(comp), _lib.ptr(conv), _lib.ptr(scratch),
                                   scratch.numel(), s), "fph_boruvka_mst")
    mst = torch.nonzero(mask[:E], as_tuple=True)[0]
    if order is not None:
        mst = order[mst].sort().values
    return mst.to(torch.int64), comp
