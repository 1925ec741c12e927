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
import scan_defs

pytestmark = pytest.mark.gpu


_colex3 = scan_defs.colex3
_numpy_scan = scan_defs.h1_definition


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
                                       cx.adj_rank, 0, complete=True)
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


_triton_walk = scan_defs.h1_reference_walk


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
                              f.vert_adj_rank, 0, complete=True)
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


def _pipeline_triangles(cx):
    """(T, 4) rows (a, b, c, largest edge rank) of the pipeline's triangles,
    in its triangle order."""
    T = cx.T
    te = cx.tri_e[:T].cpu().numpy().astype(np.int64)
    tw = cx.tri_w[:T].cpu().numpy().astype(np.int64)
    ei, ej = cx.edge_i.cpu().numpy().astype(np.int64), cx.edge_j.cpu().numpy().astype(np.int64)
    v = np.sort(np.stack([ei[te], ej[te], tw], 1), axis=1)
    return np.concatenate([v, te[:, None]], 1)


@pytest.mark.parametrize("name,pts,t", CASES, ids=[c[0] for c in CASES])
def test_scan_h2_definition_and_pipeline(name, pts, t):
    """H2 scan (reference kernels/apparent_pair_h2_kernel.py:161-317):
    flags and tetrahedron colex equal the definition on sampled triangles,
    the bounded and complete walks agree (degrees <= 2048), CPU tensors give
    the same result on the host, and on the triangles the H1 pivots leave
    the scan's set lies inside the pipeline's H2 apparent set with the same
    pivot tetrahedra."""
    from flash_ph import _pipeline
    from flash_ph.kernels.apparent_pair_h2_kernel import h2_apparent_pair_scan
    print(f"[gpu] h2 apparent scan {name}", flush=True)
    dg, cx = _pipeline.compute(torch.from_numpy(pts).cuda(), t, 2, semantics="exact")
    E, T, Q = cx.E, cx.T, cx.Q
    tri = _pipeline_triangles(cx)
    cols = [torch.from_numpy(np.ascontiguousarray(tri[:, q]).astype(np.int32)).cuda()
            for q in range(4)]
    deg = int((cx.adj_ptr[1:] - cx.adj_ptr[:-1]).max().item())
    assert deg <= 2048
    flag, colex = h2_apparent_pair_scan(*cols, cx.adj_ptr, cx.adj_idx, cx.adj_rank, deg)
    flag0, colex0 = h2_apparent_pair_scan(*cols, cx.adj_ptr, cx.adj_idx, cx.adj_rank, 0,
                                          complete=True)
    f, c = flag.cpu().numpy(), colex.cpu().numpy()
    assert np.array_equal(f, flag0.cpu().numpy()) and np.array_equal(c, colex0.cpu().numpy())
    hf, hc = h2_apparent_pair_scan(*(x.cpu() for x in cols), cx.adj_ptr.cpu(),
                                   cx.adj_idx.cpu(), cx.adj_rank.cpu(), deg)
    assert np.array_equal(hf.numpy(), f) and np.array_equal(hc.numpy(), c)
    ptr, idx, rank = cx.adj_ptr.cpu().numpy(), cx.adj_idx.cpu().numpy(), cx.adj_rank.cpu().numpy()
    sel = np.arange(T) if T <= 2500 else np.random.default_rng(0).choice(T, 2500, replace=False)
    wf, wc = scan_defs.h2_definition(tri[sel], ptr, idx, rank)
    assert np.array_equal(f[sel], wf) and np.array_equal(c[sel], wc)
    assert f.sum() > 0
    # inside the pipeline's H2 apparent set, same pivot tetrahedron
    cleared = cx.cleared[:T].cpu().numpy().astype(bool)
    kind = cx.kind2[:T].cpu().numpy()
    piv = cx.pivot2[:T].cpu().numpy().astype(np.int64)
    qe = cx.tet_e[:Q].cpu().numpy()
    key = cx.tet_key[:Q].cpu().numpy().view(np.uint64)
    ei, ej = cx.edge_i.cpu().numpy(), cx.edge_j.cpu().numpy()
    hits = np.nonzero((f == 1) & ~cleared)[0]
    assert hits.size > 0
    assert (kind[hits] == 2).all(), "scan pair not apparent in the pipeline"
    for k in hits[:3000]:
        q = int(piv[k]) - E - T
        tv = (ei[qe[q]], ej[qe[q]], int(key[q] >> np.uint64(32)),
              int(key[q] & np.uint64(0xFFFFFFFF)))
        assert scan_defs.colex4(*tv) == c[k]
    print(f"[gpu] {name}: h2 scan {int(f.sum())} of {T}; pipeline apparent "
          f"{int((kind == 2).sum())} of {int((~cleared).sum())} uncleared", flush=True)


def _hub_clusters():
    """Vertex 0 at the origin; 8000 points in 160 tight clusters of 50
    around random directions in 50 dimensions, radius 1 - 0.1 k / 8000 for
    point k (so later points are nearer the hub), cluster labels shuffled.
    t = 1.05: the hub meets every point (degree 8000 > 2048), clusters are
    cliques, distinct clusters do not meet."""
    rng = np.random.default_rng(7)
    nc, m, d = 160, 50, 50
    dirs = rng.standard_normal((nc, d))
    dirs /= np.linalg.norm(dirs, axis=1, keepdims=True)
    lab = rng.permutation(np.repeat(np.arange(nc), m))
    g = dirs[lab] + 0.09 * rng.standard_normal((nc * m, d)) / np.sqrt(d)
    g /= np.linalg.norm(g, axis=1, keepdims=True)
    r = 1.0 - 0.1 * np.arange(1, nc * m + 1) / (nc * m)
    return np.concatenate([np.zeros((1, d)), g * r[:, None]]).astype(np.float32)


def test_scan_h2_step_cap_on_a_hub():
    """Triangles (0, p, q) on the hub: the three-way walk passes the hub's
    list one vertex per step, so a witness d beyond ~6,144 is past the
    reference loop's 3 md bound (md = 2048) and the triangle is missed (its
    false negatives).  The bounded scan must equal a literal walk of that
    loop on sampled triangles, and only ever remove pairs."""
    from flash_ph.kernels.apparent_pair_h2_kernel import h2_apparent_pair_scan
    from flash_ph.rips import rips_filtration
    pts = _hub_clusters()
    print("[gpu] h2 apparent scan hub n=8001", flush=True)
    f = rips_filtration(torch.from_numpy(pts).cuda(), 1.05)
    ptr = f.vert_adj_ptr.cpu().numpy()
    idx, rank = f.vert_adj_idx.cpu().numpy(), f.vert_adj_rank.cpu().numpy()
    deg = int(np.diff(ptr).max())
    assert deg == 8000 and np.diff(ptr)[0] == 8000
    ei, ej = f.edge_i.cpu().numpy().astype(np.int64), f.edge_j.cpu().numpy().astype(np.int64)
    hub_rank = np.zeros(pts.shape[0], dtype=np.int64)
    hub_rank[idx[:ptr[1]]] = rank[:ptr[1]]
    rim = np.nonzero((ei != 0) & (ej != 0))[0]
    a = np.minimum(ei[rim], ej[rim])
    b = np.maximum(ei[rim], ej[rim])
    rt = np.maximum(rim, np.maximum(hub_rank[a], hub_rank[b]))
    tri = np.stack([np.zeros_like(a), a, b, rt], 1)
    cols = [torch.from_numpy(np.ascontiguousarray(tri[:, q]).astype(np.int32)).cuda()
            for q in range(4)]
    capped = h2_apparent_pair_scan(*cols, f.vert_adj_ptr, f.vert_adj_idx, f.vert_adj_rank, deg)
    full = h2_apparent_pair_scan(*cols, f.vert_adj_ptr, f.vert_adj_idx, f.vert_adj_rank, 0,
                                 complete=True)
    fc, cc = (x.cpu().numpy() for x in capped)
    ff, cf = (x.cpu().numpy() for x in full)
    missed = np.nonzero((ff == 1) & (fc == 0))[0]
    assert missed.size > 0                      # the bound matters here
    assert not ((fc == 1) & (ff == 0)).any()     # it only removes
    assert np.array_equal(cc[fc == 1], cf[fc == 1])
    rng = np.random.default_rng(1)
    sample = np.concatenate([missed[:60], rng.choice(len(tri), 140, replace=False)])
    for k in sample:
        a_, b_, c_, r_ = (int(x) for x in tri[k])
        d = scan_defs.h2_reference_walk(a_, b_, c_, r_, ptr, idx, rank, 2048)
        app = d >= 0 and scan_defs.h2_oldest_face(a_, b_, c_, r_, d, ptr, idx, rank)
        assert app == bool(fc[k]), k
        if app:
            assert cc[k] == scan_defs.colex4(a_, b_, c_, d)
