"""The ripser-style apparent-pair scan (reference
kernels/apparent_pair_kernel.py:101-313) on the GPU:

* flags and pivot colex equal a direct numpy evaluation of the predicate;
* its set is contained in the pipeline's forest-cleared apparent set
  (reduce_h1_gpu.py:159-223), and each flagged edge's triangle is the pivot
  the reduction assigns to that edge (apparent pairs are pairs of every
  reduction);
* with max_degree the scan reproduces the reference GPU loop's step cap on a
  vertex of degree > 2048 (a Python walk of that loop on sampled edges).
"""
from __future__ import annotations

import numpy as np
import pytest
import torch

import cases

pytestmark = pytest.mark.gpu


def _colex3(a, b, c):
    a, b, c = sorted((int(a), int(b), int(c)))
    return c * (c - 1) * (c - 2) // 6 + b * (b - 1) // 2 + a


def _numpy_scan(ei, ej, ranks, ptr, idx, rank):
    flag = np.zeros(len(ranks), dtype=np.int32)
    colex = np.zeros(len(ranks), dtype=np.int64)
    for k, re in enumerate(ranks):
        u, v = sorted((int(ei[re]), int(ej[re])))
        nu = dict(zip(idx[ptr[u]:ptr[u + 1]].tolist(), rank[ptr[u]:ptr[u + 1]].tolist()))
        best = None
        for w, rv in zip(idx[ptr[v]:ptr[v + 1]].tolist(), rank[ptr[v]:ptr[v + 1]].tolist()):
            ru = nu.get(w)
            if ru is not None and ru < re and rv < re and (best is None or w < best):
                best = w
        if best is not None:
            flag[k] = 1
            colex[k] = _colex3(u, v, best)
    return flag, colex


CASES = [("randn_d3_n300", cases.randn(300, 3, 1), 0.7),
         ("clifford_400", cases.clifford(400, 2, 0.02), 0.5),
         ("sphere_d4_n500", cases.sphere(500, 4, 3, 0.05), 0.6),
         ("randn_d10_n2000", cases.randn(2000, 10, 4), 2.2)]


@pytest.mark.parametrize("name,pts,t", CASES, ids=[c[0] for c in CASES])
def test_scan_predicate_subset_and_pivots(name, pts, t):
    from flash_ph import _pipeline
    from flash_ph.kernels.apparent_pair_kernel import apparent_pair_scan
    print(f"[gpu] apparent scan {name}", flush=True)
    dg, cx = _pipeline.compute(torch.from_numpy(pts).cuda(), t, 1, semantics="exact")
    E = cx.E
    mst = cx.mst_mask[:E].bool()
    nonmst = torch.nonzero(~mst).flatten().to(torch.int32)
    deg = int((cx.adj_ptr[1:] - cx.adj_ptr[:-1]).max().item())
    flag, colex = apparent_pair_scan(cx.edge_i, cx.edge_j, nonmst, cx.adj_ptr, cx.adj_idx,
                                     cx.adj_rank, deg)
    flag0, colex0 = apparent_pair_scan(cx.edge_i, cx.edge_j, nonmst, cx.adj_ptr, cx.adj_idx,
                                       cx.adj_rank, 0)
    torch.cuda.synchronize()
    f, c = flag.cpu().numpy(), colex.cpu().numpy()
    assert deg <= 2048
    assert np.array_equal(f, flag0.cpu().numpy()) and np.array_equal(c, colex0.cpu().numpy())
    ei, ej = cx.edge_i.cpu().numpy(), cx.edge_j.cpu().numpy()
    ptr, idx, rank = cx.adj_ptr.cpu().numpy(), cx.adj_idx.cpu().numpy(), cx.adj_rank.cpu().numpy()
    nm = nonmst.cpu().numpy()
    sel = np.arange(nm.size) if nm.size <= 4000 else np.random.default_rng(0).choice(
        nm.size, 4000, replace=False)
    wf, wc = _numpy_scan(ei, ej, nm[sel], ptr, idx, rank)
    assert np.array_equal(f[sel], wf) and np.array_equal(c[sel], wc)
    assert f.sum() > 0
    # subset of the forest-cleared set, and the reduction's pivot
    kind = cx.kind1[:E].cpu().numpy()
    piv = cx.pivot1[:E].cpu().numpy()
    tri_e, tri_w = cx.tri_e.cpu().numpy(), cx.tri_w.cpu().numpy()
    for k in np.nonzero(f)[0]:
        e = int(nm[k])
        assert kind[e] == 2, f"edge {e} flagged but not apparent in the pipeline"
        t_ = int(piv[e]) - E
        assert _colex3(ei[tri_e[t_]], ej[tri_e[t_]], tri_w[t_]) == c[k]
    print(f"[gpu] {name}: scan {int(f.sum())} of {int((kind == 2).sum())} apparent", flush=True)


def _triton_walk(u, v, re, ptr, idx, rank, md):
    """The reference's per-edge merge loop (kernels/apparent_pair_kernel.py:
    133-156) with its 2 md step bound, one edge at a time."""
    iu, iv = int(ptr[u]), int(ptr[v])
    ue, ve = int(ptr[u + 1]), int(ptr[v + 1])
    for _ in range(2 * md):
        if iu >= ue or iv >= ve:
            break
        nu, nv = idx[iu], idx[iv]
        if nu == nv:
            if rank[iu] < re and rank[iv] < re:
                return int(nu)
            iu += 1
            iv += 1
        elif nu < nv:
            iu += 1
        else:
            iv += 1
    return -1


def test_scan_step_cap_on_a_hub():
    """Vertex 0 at the centre of 6000 points at random angles, point k at
    radius 1 - 1e-5 k, t = 1.05: the hub has degree 6000 > 2048, its edges
    get shorter with k, so the witnesses of hub edge (0, p) are points after
    p, reached only after ~p merge steps; past the reference loop's
    4096-step bound they are missed (its false negatives)."""
    from flash_ph.kernels.apparent_pair_kernel import apparent_pair_scan
    from flash_ph.rips import rips_filtration
    rng = np.random.default_rng(5)
    th = rng.uniform(0, 2 * np.pi, 6000)
    r = 1.0 - np.arange(1, 6001) * 1e-5
    pts = np.concatenate([np.zeros((1, 2)), np.column_stack([r * np.cos(th), r * np.sin(th)])])
    pts = pts.astype(np.float32)
    print("[gpu] apparent scan hub n=6001", flush=True)
    f = rips_filtration(torch.from_numpy(pts).cuda(), 1.05)
    ptr = f.vert_adj_ptr.cpu().numpy()
    deg = int(np.diff(ptr).max())
    assert deg == 6000
    ei, ej = f.edge_i.cpu().numpy(), f.edge_j.cpu().numpy()
    hub = np.nonzero(ei == 0)[0].astype(np.int32)
    ranks = torch.from_numpy(hub).cuda()
    capped = apparent_pair_scan(f.edge_i, f.edge_j, ranks, f.vert_adj_ptr, f.vert_adj_idx,
                                f.vert_adj_rank, deg)
    full = apparent_pair_scan(f.edge_i, f.edge_j, ranks, f.vert_adj_ptr, f.vert_adj_idx,
                              f.vert_adj_rank, 0)
    fc, cc = (x.cpu().numpy() for x in capped)
    ff, cf = (x.cpu().numpy() for x in full)
    missed = np.nonzero((ff == 1) & (fc == 0))[0]
    assert missed.size > 0                      # the cap matters here
    assert not ((fc == 1) & (ff == 0)).any()     # it only removes
    idx, rank = f.vert_adj_idx.cpu().numpy(), f.vert_adj_rank.cpu().numpy()
    rng2 = np.random.default_rng(1)
    sample = np.concatenate([missed[:40], rng2.choice(hub.size, 60, replace=False)])
    for k in sample:
        e = int(hub[k])
        u, v = sorted((int(ei[e]), int(ej[e])))
        w = _triton_walk(u, v, e, ptr, idx, rank, 2048)
        assert (w >= 0) == bool(fc[k]), k
        if w >= 0:
            assert cc[k] == _colex3(u, v, w)
