"""Public API for Vietoris-Rips persistent homology (H0 + H1 + H2).

Mirrors the reference persistence.py.  Pipeline, all on one GPU:
edge screening -> filtration sort and CSR -> Boruvka forest (H0) ->
triangles, apparent pairs and residual reduction (H1) -> H1->H2 clearing,
tetrahedra, apparent pairs and residual reduction (H2).
"""
from __future__ import annotations

import numpy as np
import torch
from torch import Tensor

from flash_ph import _lib, _pipeline

_VERTEX_PACK_BITS = 16
_VALID_BACKENDS = ("auto", "gpu", "giotto")
# accepted in addition to the reference's backends
_EXTRA_BACKENDS = ("exact",)


def _rips_h1h2_giotto(filt, n, max_dim, max_edge_length, device):
    """H1 (and H2) from giotto-ph's reducer, for ``backend='giotto'``
    (the reference's route, persistence.py:27-61).  The filtration leaves the
    GPU as COO triplets written by fph_coo_export (one device-to-host copy);
    giotto-ph then reduces with edge collapse on all host cores.  Raises
    ModuleNotFoundError when giotto-ph is not installed."""
    from gph import ripser_parallel

    dm = coo_distance_matrix(filt)
    res = ripser_parallel(dm, maxdim=max_dim, thresh=max_edge_length, metric="precomputed",
                          collapse_edges=True, n_threads=-1)
    out = []
    for dim in range(1, max_dim + 1):
        rows = np.asarray(res["dgms"][dim], dtype=np.float32).reshape(-1, 2)
        t = torch.from_numpy(rows).to(device)
        out.append(t[torch.argsort(t[:, 0], stable=True)] if t.shape[0] else
                   torch.empty(0, 2, dtype=torch.float32, device=device))
    return out


def _rips_persistence_cpu(points, max_edge_length, max_dim, backend, exact_colex):
    """Diagrams of a cloud in host memory with libflashph_cpu (OpenMP).  H1
    and H2 are complete reductions, the pairs giotto-ph returns on the
    reference's CPU route (persistence.py:197-206); backend='giotto' hands
    them to giotto-ph itself."""
    from flash_ph import _cpu
    device = points.device
    pts = points.detach().float().contiguous().numpy()
    thr = _pipeline.threshold_sq_f32(max_edge_length)
    if max_dim >= 1 and backend == "giotto":
        out = _cpu.rips(pts, thr, 0, exact_colex=exact_colex)
        f = _cpu.filtration(pts, thr, exact_colex=exact_colex)
        filt = _HostFiltration(f, pts.shape[0])
        h = _rips_h1h2_giotto(filt, pts.shape[0], max_dim, max_edge_length, device)
        return [torch.from_numpy(out[0]).to(device)] + h
    dg = _cpu.rips(pts, thr, max_dim, exact_h1=True, exact_colex=exact_colex)
    return [torch.from_numpy(d).to(device) for d in dg]


class _HostFiltration:
    """The fields coo_distance_matrix reads, from the host filtration."""

    def __init__(self, f, n):
        self.n = n
        self.edge_i = torch.from_numpy(f["edge_i"])
        self.edge_j = torch.from_numpy(f["edge_j"])
        self.edge_dist_sq = torch.from_numpy(f["edge_dsq"])


def coo_distance_matrix(filt):
    """scipy (n, n) COO sparse distance matrix of a RipsFiltration: one entry
    (i, j, |x_i - x_j|) per edge (i < j; FP32 correctly rounded sqrt of the
    edge's dsq) and the zero diagonal -- the input the reference hands to
    giotto-ph (persistence.py:36-45)."""
    import scipy.sparse

    n = int(filt.n)
    E = int(filt.edge_i.shape[0])
    dev = filt.edge_i.device
    if dev.type != "cuda":   # host filtration: the same triplets with numpy
        i = filt.edge_i.numpy().astype(np.int64)
        j = filt.edge_j.numpy().astype(np.int64)
        v = np.sqrt(filt.edge_dist_sq.numpy().astype(np.float64)).astype(np.float32)
        diag = np.arange(n, dtype=np.int64)
        return scipy.sparse.coo_matrix(
            (np.concatenate([v, np.zeros(n, np.float32)]),
             (np.concatenate([i, diag]), np.concatenate([j, diag]))), shape=(n, n))
    lib = _lib.load()
    row = torch.empty(E + n, dtype=torch.int64, device=dev)
    col = torch.empty(E + n, dtype=torch.int64, device=dev)
    val = torch.empty(E + n, dtype=torch.float32, device=dev)
    _lib.check(lib.fph_coo_export(_lib.ptr(filt.edge_i), _lib.ptr(filt.edge_j),
                                  _lib.ptr(filt.edge_dist_sq), E, n, row.data_ptr(), col.data_ptr(),
                                  val.data_ptr(), _lib.stream_of(dev)), "fph_coo_export")
    host = torch.cat([row, col, val.view(torch.int32).long()]).cpu().numpy()
    r, c = host[:E + n], host[E + n:2 * (E + n)]
    v = host[2 * (E + n):].astype(np.int32).view(np.float32)
    return scipy.sparse.coo_matrix((v, (r, c)), shape=(n, n))


def rips_persistence(
    points,
    max_edge_length: float,
    max_dim: int = 1,
    backend: str = 'auto',
    *,
    large_h2: bool = False,
    exact_colex: bool = False,
) -> list[Tensor]:
    """Exact Vietoris-Rips persistent homology (H0 + H1 + optional H2).

    Parameters
    ----------
    points : (n, d) array-like
        Point cloud coordinates (any d >= 1).  NumPy arrays are converted to
        a float32 tensor on the GPU when one is visible.  CUDA tensors run on
        the GPU (gfx950 HIP library); CPU tensors run on the host pipeline
        (C++/OpenMP, the same arithmetic: bit-identical diagrams), as the
        reference computes CPU tensors on the CPU.
    max_edge_length : float
        Maximum edge length (strict < inequality).
    max_dim : int
        0 (H0 only), 1 (H0 + H1), or 2 (H0 + H1 + H2).  Default 1.
        H2 uses 16-bit vertex packing, requiring n < 65536.
    backend : str
        H1 reduction backend (validated when ``max_dim=1``).  ``'auto'`` and
        ``'gpu'`` reduce on the GPU with the reference's own semantics,
        including the capacity of its column store (a reduced column that
        does not fit is not stored, reduce_h1_gpu.py:282-289), so the pairs
        equal the reference's in every case; ``'exact'`` runs the parallel
        complete reduction (equal to ``'gpu'`` whenever the reference's store
        does not overflow); ``'giotto'`` hands H1 to giotto-ph.  For
        ``max_dim=2`` H1 and H2 are complete reductions, as giotto-ph's.

    large_h2 : bool, keyword only
        Allow ``max_dim=2`` for n >= 65536.  The reference packs tetrahedron
        vertices in 16 bits and raises ValueError there
        (persistence.py:125-129), which stays the default; this build packs
        them in 32 bits.
    exact_colex : bool, keyword only
        For n > 92,682, break ties between equal edge lengths by the full
        colex index j(j-1)/2 + i instead of the reference's 32-bit
        truncation (rips.py:36-49) and emission order.  Changes the row
        order within equal births, never the intervals.

    Returns
    -------
    list[Tensor]
        [H0, H1, ...], each (K, 2) float32 on the input's device: H0 has one
        (0, inf) row per connected component; H1/H2 include essential
        (birth, inf) rows; rows sorted stably by birth.  When ``points``
        requires grad, the diagrams carry gradients back to the points
        (each finite value is an edge length; see flash_ph.autograd).

    Reference: persistence.py:68-206.
    """
    if isinstance(points, np.ndarray):
        points = torch.from_numpy(points).float()
        if torch.cuda.is_available():
            points = points.cuda()

    if max_edge_length < 0:
        raise ValueError(
            f"max_edge_length must be non-negative, got {max_edge_length}"
        )
    if points.ndim != 2:
        raise ValueError(
            f"points must have shape (n, d), got {tuple(points.shape)}"
        )
    if max_dim not in (0, 1, 2):
        raise ValueError(f"max_dim must be 0, 1, or 2, got {max_dim}")
    if max_dim == 2 and points.shape[0] >= (1 << _VERTEX_PACK_BITS) and not large_h2:
        raise ValueError(
            f"H2 requires n < {1 << _VERTEX_PACK_BITS} (16-bit vertex packing), "
            f"got n={points.shape[0]}. Use max_dim=1 or reduce n."
        )

    device = points.device
    n = points.shape[0]

    if n == 0:
        return [
            torch.empty(0, 2, dtype=torch.float32, device=device)
            for _ in range(max_dim + 1)
        ]
    if n == 1:
        h0 = torch.tensor([[0.0, float('inf')]], dtype=torch.float32, device=device)
        return [h0] + [
            torch.empty(0, 2, dtype=torch.float32, device=device)
            for _ in range(max_dim)
        ]

    if max_dim == 1 and backend not in _VALID_BACKENDS + _EXTRA_BACKENDS:
        # reduce_h1_gpu.py:462-465
        raise ValueError(
            f"backend must be one of {_VALID_BACKENDS}, got {backend!r}"
        )

    if device.type != "cuda":
        # host memory: the host pipeline, as the reference computes CPU
        # tensors on the CPU (rips.py:221-233, persistence.py:197-206)
        return _rips_persistence_cpu(points, max_edge_length, max_dim, backend, exact_colex)

    _lib.require_gpu()
    pts = points.float().contiguous()

    if torch.is_grad_enabled() and points.requires_grad and backend != "giotto":
        from flash_ph.autograd import rips_persistence_diff
        semantics = "exact" if backend == "exact" else "reference"
        out = rips_persistence_diff(pts, max_edge_length, max_dim, semantics)
        return out if device.type == "cuda" else [d.to(device) for d in out]

    if max_dim == 1 and backend == "giotto":
        from flash_ph.rips import RipsFiltration
        cx = _pipeline.build_edges(pts, max_edge_length)
        _pipeline.build_mst(cx)
        filt = RipsFiltration(cx.edge_i, cx.edge_j, cx.edge_dsq, cx.adj_ptr, cx.adj_idx,
                              cx.adj_rank, cx.adj_dsq, n)
        h1 = _rips_h1h2_giotto(filt, n, 1, max_edge_length, pts.device)[0]
        out = [cx.h0, h1]
    elif max_dim == 0 and n <= _pipeline.MAX_KEY_DECODE_N:
        # H0 needs only the forest: Boruvka on the screened keys, no edge
        # sort or CSR (and no 2^31 edge limit)
        out = [_pipeline.h0_direct(pts, max_edge_length)]
    else:
        semantics = "exact" if backend == "exact" else "reference"
        out, _ = _pipeline.compute(pts, max_edge_length, max_dim, semantics=semantics,
                                   colex64=exact_colex)
    if device.type != "cuda" or out[0].device != device:
        out = [d.to(device) for d in out]
    return out
