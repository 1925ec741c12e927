"""The reference's size limits, lifted (GPU):

* H0 from the screened keys (fph_h0_from_keys): the oracle's H0 on the
  parity families, and a cloud with more than 2^31 edges -- 66,000 integer
  points on a line, every pair an edge (2,177,967,000 edges, past the
  reference's E < 2^31 assertion, kernels/edge_kernel.py:203-205) -- whose
  minimum spanning forest is the path through the sorted points, so H0 is
  the sorted gaps;
* H2 for n >= 65,536 (large_h2=True; the reference raises there,
  persistence.py:125-129): a sparse 70,000-point cloud against the oracle
  run from the same edge set, the edge set checked row by row on sampled
  vertices.
"""
from __future__ import annotations

import numpy as np
import pytest
import torch

import cases

pytestmark = pytest.mark.gpu


def _u32(a):
    return np.ascontiguousarray(a).view(np.uint32)


H0_CASES = [("randn_d3_n500", cases.randn(500, 3, 1), 0.6),
            ("clifford_1500", cases.clifford(1500, 2, 0.02), 0.3),
            ("dup_points", np.repeat(cases.randn(200, 2, 3), 3, axis=0), 0.2),
            ("no_edges", cases.randn(50, 4, 4), 1e-3),
            ("two_points", cases.randn(2, 3, 5), 10.0),
            ("randn_d50_n5000_mfma", cases.randn(5000, 50, 6), 8.6),
            ("randn_d10_n20000", cases.randn(20_000, 10, 7), 1.2)]


@pytest.mark.parametrize("name,pts,t", H0_CASES, ids=[c[0] for c in H0_CASES])
def test_h0_from_keys_matches_oracle(name, pts, t):
    import oracle
    from flash_ph import _pipeline, rips_persistence
    print(f"[gpu] h0 from keys {name}", flush=True)
    x = torch.from_numpy(np.ascontiguousarray(pts, dtype=np.float32)).cuda()
    h0 = rips_persistence(x, t, max_dim=0)[0].cpu().numpy()
    ref = oracle.run(pts, t, 0)["h0"]
    assert h0.shape == ref.shape and np.array_equal(_u32(h0), _u32(ref))
    # and equal to the sorted-edge path (edges, CSR, rank Boruvka)
    cx = _pipeline.build_edges(x, t)
    _pipeline.build_mst(cx)
    assert np.array_equal(_u32(cx.h0.cpu().numpy()), _u32(h0))


def test_h0_beyond_2pow31_edges():
    from flash_ph import rips_persistence
    n = 66_000
    rng = np.random.default_rng(11)
    gaps = rng.integers(1, 4, n - 1)
    xs = np.concatenate([[0], np.cumsum(gaps)]).astype(np.float32)  # exact integers
    pts = xs[rng.permutation(n)][:, None].copy()
    t = 200_000.0        # every pair: the longest span is < 2 * 66,000
    E = n * (n - 1) // 2
    assert E > 2**31
    print(f"[gpu] h0 with {E} edges", flush=True)
    h0 = rips_persistence(torch.from_numpy(pts).cuda(), t, max_dim=0)[0].cpu().numpy()
    want = np.empty((n, 2), dtype=np.float32)
    want[:, 0] = 0.0
    want[:-1, 1] = np.sort(gaps.astype(np.float32))   # sqrt(g^2) = g exactly
    want[-1, 1] = np.inf
    assert np.array_equal(_u32(h0), _u32(want))


def test_h2_beyond_65536_points():
    import oracle
    from flash_ph import _pipeline, rips_persistence
    n, t = 70_000, 1.2
    rng = np.random.default_rng(12)
    pts = (rng.random((n, 3)) * 40.0).astype(np.float32)
    with pytest.raises(ValueError, match="H2 requires n < 65536"):
        rips_persistence(torch.from_numpy(pts), t, max_dim=2)
    x = torch.from_numpy(pts).cuda()
    print(f"[gpu] h2 n={n}", flush=True)
    dg = rips_persistence(x, t, max_dim=2, large_h2=True)
    cx = _pipeline.build_edges(x, t)
    ei, ej, ed = (a.cpu().numpy() for a in (cx.edge_i, cx.edge_j, cx.edge_dsq))
    ptr, idx = cx.adj_ptr.cpu().numpy(), cx.adj_idx.cpu().numpy()
    for v in rng.choice(n, 24, replace=False):
        oi, _ = oracle.vertex_neighbours(pts, t, int(v))
        assert np.array_equal(idx[ptr[v]:ptr[v + 1]], oi)
    print("[gpu] oracle from the edge set", flush=True)
    ref = oracle.run(pts, t, 2, edges=(ei, ej, ed))
    assert ref["stats"]["Q"] > 0 and (ref["edge_j"] >= 65536).any()
    for k in range(3):
        got = dg[k].cpu().numpy()
        want = ref["diagrams"][k]
        assert got.shape == want.shape and np.array_equal(_u32(got), _u32(want)), f"H{k}"


def test_exact_colex_order_and_invariance():
    """exact_colex=True (n > 92,682): edges in (dsq, j, i) order; the
    diagrams' intervals are those of the reference's order (a filtration's
    diagram does not depend on how equal values are ordered), which itself
    matches the oracle bit for bit, row order included."""
    import oracle
    from flash_ph import _pipeline, rips_persistence
    pts = cases.grid_points(47)
    print("[gpu] exact colex order on the grid", flush=True)
    cx = _pipeline.build_edges(torch.from_numpy(pts).cuda(), 1.5, colex64=True)
    i, j, dsq = cases.grid_edges(pts, 1.5)
    order = np.lexsort((i, j, dsq.view(np.uint32)))
    assert np.array_equal(cx.edge_i.cpu().numpy(), i[order])
    assert np.array_equal(cx.edge_j.cpu().numpy(), j[order])
    assert not np.array_equal(order, cases.reference_edge_order(i, j, dsq))
    rng = np.random.default_rng(13)
    flat = rng.choice(1000 * 1000, 93_000, replace=False)
    pts2 = np.column_stack([flat // 1000, flat % 1000]).astype(np.float32)
    x = torch.from_numpy(pts2).cuda()
    print("[gpu] exact colex invariance n=93000", flush=True)
    ref_dg = rips_persistence(x, 2.9, max_dim=1)
    ex_dg = rips_persistence(x, 2.9, max_dim=1, exact_colex=True)
    c2 = _pipeline.build_edges(x, 2.9)
    ei, ej, ed = (a.cpu().numpy() for a in (c2.edge_i, c2.edge_j, c2.edge_dsq))
    ref = oracle.run(pts2, 2.9, 1, edges=(ei, ej, ed))
    for k in range(2):
        a, b = ref_dg[k].cpu().numpy(), ex_dg[k].cpu().numpy()
        assert np.array_equal(_u32(a), _u32(ref["diagrams"][k])), f"H{k} vs oracle"
        key = lambda r: r[np.lexsort((r[:, 1], r[:, 0]))]
        assert np.array_equal(_u32(key(a)), _u32(key(b))), f"H{k} multiset"
