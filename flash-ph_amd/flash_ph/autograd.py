"""Differentiable persistence diagrams.

Every finite value of a Vietoris-Rips diagram is the length of one edge:
the longest edge (diameter) of the simplex that creates or destroys the
class.  The diagram is therefore a piecewise-smooth function of the points,
with gradient
    d |x_u - x_v| / d x_u = (x_u - x_v) / |x_u - x_v|
scattered to the two endpoints of each critical edge.  ``forward`` runs the
pipeline once, recording the critical edges of every row (fph_diagram's
out_edges, and the forest edges for H0); ``backward`` hands the upstream
gradients and those edges to fph_backward_scatter, a deterministic HIP
kernel.  H0 births are constant 0 and infinite deaths carry no gradient.

The reference has no gradient path; ``rips_persistence`` and
``rips_persistence_batch`` switch to these functions when an input
requires grad, and return the same values.
"""
from __future__ import annotations

import torch

from flash_ph import _lib, _pipeline


def _layout_tensors(clouds, device):
    """foff, dims, vbase of clouds (device tensors) plus host vertex offsets."""
    ns = [int(c.shape[0]) for c in clouds]
    ds = [int(c.shape[1]) for c in clouds]
    foff, vbase = [0], [0]
    for n, d in zip(ns, ds):
        foff.append(foff[-1] + n * d)
        vbase.append(vbase[-1] + n)
    i64 = torch.tensor(foff + vbase, dtype=torch.int64).to(device)
    k = len(clouds) + 1
    dims = torch.tensor(ds, dtype=torch.int32).to(device)
    return i64[:k], dims, i64[k:2 * k], vbase


def _h0_edges(mst_mask, E, edge_off, vbase, N, device):
    """(N, 2) int32: per H0 row, -1 for the birth and the forest edge of the
    death (-1 for the essential rows).  Cloud c's rows start at vbase[c]; its
    finite rows are its forest edges in filtration order."""
    out = torch.full((max(N, 1), 2), -1, dtype=torch.int32, device=device)
    if E == 0 or N == 0:
        return out[:N]
    fe = torch.nonzero(mst_mask[:E]).flatten()            # ascending, grouped by cloud
    if fe.numel() == 0:
        return out[:N]
    eo = edge_off.long()
    cloud = torch.searchsorted(eo, fe, right=True) - 1
    first = torch.searchsorted(fe, eo[:-1])               # first forest edge of each cloud
    vb = torch.tensor(vbase[:-1], dtype=torch.int64, device=device)
    row = vb[cloud] + (torch.arange(fe.numel(), device=device) - first[cloud])
    out[row, 1] = fe.to(torch.int32)
    return out[:N]


def _scatter_grad(ctx, grads_per_dim):
    """grads_per_dim: list over (diagram, row_edges) of upstream gradients
    (None where unused).  Returns the gradient of the packed points."""
    ent_e, ent_v = [], []
    for g, re in grads_per_dim:
        if g is None or re is None or re.shape[0] == 0:
            continue
        ent_e.append(re.reshape(-1))
        ent_v.append(g.detach().to(torch.float32).reshape(-1).contiguous())
    dev = ctx.pts.device
    if ent_e:
        edges = torch.cat(ent_e).to(torch.int32).contiguous()
        vals = torch.cat(ent_v).contiguous()
        # an infinite value carries no gradient (and a NaN upstream gradient
        # on it must not leak in)
        vals = torch.where(edges >= 0, vals, torch.zeros_like(vals))
    else:
        edges = torch.empty(0, dtype=torch.int32, device=dev)
        vals = torch.empty(0, dtype=torch.float32, device=dev)
    cx = ctx.cx
    grad = torch.empty_like(ctx.pts)
    _lib.call_ws("fph_backward_scatter",
                 (_lib.ptr(ctx.pts), _lib.ptr(ctx.foff), _lib.ptr(ctx.dims),
                  _lib.ptr(ctx.vbase), ctx.C, ctx.N, _lib.ptr(cx.edge_filt), cx.E,
                  _lib.ptr(cx.adj_ptr), _lib.ptr(cx.adj_idx), _lib.ptr(cx.adj_rank),
                  _lib.ptr(edges), _lib.ptr(vals), edges.numel(), _lib.ptr(grad)),
                 (_lib.stream_of(dev),), dev)
    return grad


class RipsPersistenceFunction(torch.autograd.Function):
    """One cloud: returns (H0, H1, ...) with gradients to the points."""

    @staticmethod
    def forward(ctx, points, max_edge_length, max_dim, semantics):
        pts = points.detach().float().contiguous()
        n = pts.shape[0]
        dgms, cx = _pipeline.compute(pts, max_edge_length, max_dim, semantics=semantics,
                                     want_edges=True)
        ctx.pts, ctx.cx, ctx.C, ctx.N = pts.reshape(-1), cx, 1, n
        ctx.foff, ctx.dims, ctx.vbase, vb = _layout_tensors([pts], pts.device)
        eo = torch.tensor([0, cx.E], dtype=torch.int32, device=pts.device)
        ctx.row_edges = [_h0_edges(cx.mst_mask, cx.E, eo, vb, n, pts.device)]
        if max_dim >= 1:
            ctx.row_edges.append(cx.h1_edges)
        if max_dim >= 2:
            ctx.row_edges.append(cx.h2_edges)
        ctx.in_dtype = points.dtype
        return tuple(d.clone() for d in dgms)

    @staticmethod
    def backward(ctx, *grads):
        g = _scatter_grad(ctx, list(zip(grads, ctx.row_edges)))
        return g.reshape(ctx.N, -1).to(ctx.in_dtype), None, None, None


class RipsPersistenceBatchFunction(torch.autograd.Function):
    """A batch of clouds: returns H0, H1 of cloud 0, then of cloud 1, ...
    with gradients to every cloud."""

    @staticmethod
    def forward(ctx, spec, *clouds):
        thr_sq, max_dim, semantics = spec
        pts = [c.detach().float().contiguous() for c in clouds]
        dev = pts[0].device
        L = _pipeline.layout_batch(pts, dev)
        dgms, cx = _pipeline.compute_batch(L, thr_sq, max_dim, semantics=semantics,
                                           want_edges=True)
        ctx.ns = [int(c.shape[0]) for c in clouds]
        ctx.ds = [int(c.shape[1]) for c in clouds]
        ctx.dtypes = [c.dtype for c in clouds]
        ctx.pts, ctx.cx, ctx.C, ctx.N = L.pts, cx, L.C, L.N
        ctx.foff, ctx.dims, ctx.vbase = L.foff, L.dims, L.vbase
        ctx.max_dim = max_dim
        ctx.row_edges = []
        if cx is not None:
            h0e = _h0_edges(cx.mst_mask, cx.E, cx.edge_off, L.vbase_host, L.N, dev)
            vb = L.vbase_host
            for c in range(L.C):
                ctx.row_edges.append(h0e[vb[c]:vb[c + 1]])
                if max_dim >= 1:
                    ctx.row_edges.append(cx.h1_edges[c])
        return tuple(d.clone() for per in dgms for d in per)

    @staticmethod
    def backward(ctx, *grads):
        if ctx.cx is None:
            return (None,) + tuple(torch.zeros(n, d, dtype=t, device=ctx.pts.device)
                                   for n, d, t in zip(ctx.ns, ctx.ds, ctx.dtypes))
        g = _scatter_grad(ctx, list(zip(grads, ctx.row_edges)))
        out, at = [], 0
        for n, d, t in zip(ctx.ns, ctx.ds, ctx.dtypes):
            out.append(g[at:at + n * d].reshape(n, d).to(t))
            at += n * d
        return (None,) + tuple(out)


def rips_persistence_diff(points, max_edge_length, max_dim=1, semantics="reference"):
    return list(RipsPersistenceFunction.apply(points, max_edge_length, max_dim, semantics))


def rips_persistence_batch_diff(clouds, thr_sq, max_dim, semantics):
    flat = RipsPersistenceBatchFunction.apply((thr_sq, max_dim, semantics), *clouds)
    k = max_dim + 1
    return [list(flat[i:i + k]) for i in range(0, len(flat), k)]
