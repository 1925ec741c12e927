"""Pins the CPU oracle (oracle/fph_oracle.c) with checks that do not share
its code: exact-rational FMA, numpy lexsort, scipy's spanning tree, brute
clique enumeration, brute-force homology reduction, and the reference's own
known-answer tests (testing/test_persistence.py, testing/test_reduce_h1.py)."""
from __future__ import annotations

import itertools

import numpy as np
import pytest
from scipy.sparse import coo_matrix
from scipy.sparse.csgraph import minimum_spanning_tree

import brute
import cases
import oracle


def _bits_equal(a, b):
    a = np.ascontiguousarray(a, np.float32)
    b = np.ascontiguousarray(b, np.float32)
    return a.shape == b.shape and np.array_equal(a.view(np.uint32), b.view(np.uint32))


# ---------------------------------------------------------------- numerics

def test_fmaf_matches_exact_rounding():
    rng = np.random.default_rng(1)
    vals = rng.standard_normal((3000, 3)).astype(np.float32)
    # products whose double rounding would differ from a single rounding
    vals[:200, 2] = np.float32(1.0)
    vals[:200, 0] = (np.float32(1.0) + np.float32(2.0 ** -12)).astype(np.float32)
    vals[:200, 1] = vals[:200, 0]
    for a, b, c in vals:
        assert _bits_equal(oracle.fmaf(a, b, c), brute.fmaf_exact(a, b, c))


@pytest.mark.parametrize("d", [1, 3, 10, 50])
def test_distance_chain_bit_exact(d):
    pts = cases.randn(12, d, 100 + d)
    for i, j in itertools.combinations(range(12), 2):
        assert _bits_equal(oracle.dist_sq(pts, i, j), brute.fma_chain_exact(pts[i], pts[j]))


def test_threshold_square_rounding():
    # t*t in double then FP32 (rips.py:125): 1.6 -> 2.559999942779541
    assert float(oracle.threshold_sq_f32(1.6)) == np.float32(2.559999942779541)


def test_edges_against_fp64():
    pts = cases.randn(400, 10, 9)
    t = 2.4
    r = oracle.run(pts, t, 0)
    thr = oracle.threshold_sq_f32(t)
    d64 = ((pts[:, None, :].astype(np.float64) - pts[None, :, :]) ** 2).sum(-1)
    iu, ju = np.triu_indices(400, 1)
    ref_in = d64[iu, ju] < float(thr)
    got = set(zip(r["edge_i"].tolist(), r["edge_j"].tolist()))
    band = np.abs(d64[iu, ju] - float(thr)) <= 1e-5 * float(thr)
    for k in np.nonzero(~band)[0]:
        assert ((iu[k], ju[k]) in got) == bool(ref_in[k])
    # FP32 chain within a few ulp of the FP64 value
    e64 = d64[r["edge_i"], r["edge_j"]]
    rel = np.abs(r["edge_dsq"].astype(np.float64) - e64) / np.maximum(e64, 1e-30)
    assert rel.max() < 1e-5


# ---------------------------------------------------- order, CSR, forest, H0

@pytest.mark.parametrize("name,make,t,md", cases.SMALL[:8], ids=[c[0] for c in cases.SMALL[:8]])
def test_sort_and_csr_closed_form(name, make, t, md):
    pts = make()
    r = oracle.run(pts, t, 0)
    n = pts.shape[0]
    ei, ej, dq = r["edge_i"], r["edge_j"], r["edge_dsq"]
    assert np.all(ei < ej)
    colex = ej.astype(np.int64) * (ej.astype(np.int64) - 1) // 2 + ei
    order = np.lexsort((colex, dq.view(np.uint32)))
    assert np.array_equal(order, np.arange(len(ei)))
    src = np.concatenate([ei, ej]); dst = np.concatenate([ej, ei])
    rank = np.concatenate([np.arange(len(ei))] * 2)
    o = np.lexsort((dst, src))
    assert np.array_equal(r["adj_idx"], dst[o])
    assert np.array_equal(r["adj_rank"], rank[o])
    assert np.array_equal(r["adj_ptr"], np.concatenate([[0], np.cumsum(np.bincount(src, minlength=n))]))


@pytest.mark.parametrize("name,make,t,md", cases.SMALL + cases.MEDIUM[:2],
                         ids=[c[0] for c in cases.SMALL + cases.MEDIUM[:2]])
def test_forest_matches_scipy(name, make, t, md):
    pts = make()
    r = oracle.run(pts, t, 0)
    n, E = pts.shape[0], len(r["edge_i"])
    assert np.array_equal(r["mst_idx"], np.nonzero(r["mst_kruskal"])[0])
    if E:
        g = coo_matrix((np.arange(1, E + 1, dtype=np.float64), (r["edge_i"], r["edge_j"])),
                       shape=(n, n)).tocsr()
        tree = minimum_spanning_tree(g).tocoo()
        assert np.array_equal(np.sort(tree.data.astype(np.int64) - 1), r["mst_idx"])
    h0 = r["h0"]
    assert h0.shape == (n, 2)
    fin = h0[np.isfinite(h0[:, 1]), 1]
    assert _bits_equal(fin, np.sqrt(r["edge_dsq"][r["mst_idx"]]))
    assert int(np.isinf(h0[:, 1]).sum()) == n - len(r["mst_idx"])


# ------------------------------------------------------- simplices, pairs

def _tiny_cases():
    return [
        ("randn_d2_n18", cases.randn(18, 2, 1), 1.2),
        ("randn_d3_n16", cases.randn(16, 3, 2), 1.6),
        ("randn_d4_n14", cases.randn(14, 4, 3), 2.2),
        ("randn_d10_n12", cases.randn(12, 10, 4), 4.0),
        ("circle_12", cases.circle(12), 1.2),
        ("sphere_14", cases.sphere(14, 3, 6, 0.05), 1.5),
        ("grid_2x2x2", cases.grid3()[[0, 1, 3, 4, 9, 10, 12, 13]], 1.5),
        ("octahedron", np.array([[1, 0, 0], [-1, 0, 0], [0, 1, 0], [0, -1, 0], [0, 0, 1],
                                 [0, 0, -1]], np.float32), 1.5),
        ("dupes", np.repeat(cases.randn(6, 2, 5), 2, axis=0), 1.5),
    ]


@pytest.mark.parametrize("name,pts,t", _tiny_cases(), ids=[c[0] for c in _tiny_cases()])
def test_simplices_match_clique_enumeration(name, pts, t):
    r = oracle.run(pts, t, 2)
    thr = oracle.threshold_sq_f32(t)
    D = brute.all_dsq(pts)
    simp = brute.rips_cliques(D, np.float32(thr), 3)
    tri = {s for s, _ in simp if len(s) == 3}
    tet = {s for s, _ in simp if len(s) == 4}
    got_tri = set(zip(r["tri_v0"].tolist(), r["tri_v1"].tolist(), r["tri_v2"].tolist()))
    got_tet = set(zip(r["tet_v0"].tolist(), r["tet_v1"].tolist(), r["tet_v2"].tolist(),
                      r["tet_v3"].tolist()))
    assert got_tri == tri
    assert got_tet == tet


@pytest.mark.parametrize("name,pts,t", _tiny_cases(), ids=[c[0] for c in _tiny_cases()])
def test_diagrams_match_brute_homology(name, pts, t):
    r = oracle.run(pts, t, 2)
    assert r["stats"]["pool_drops"] == 0 and r["stats"]["truncations"] == 0
    ref = brute.brute_diagrams(pts, oracle.threshold_sq_f32(t), 2)
    for k in range(3):
        assert _bits_equal(brute.multiset(r["diagrams"][k]), brute.multiset(ref[k])), f"H{k}"


@pytest.mark.parametrize("seed", range(12))
def test_random_diagrams_match_brute_homology(seed):
    rng = np.random.default_rng(1000 + seed)
    n = int(rng.integers(8, 17))
    d = int(rng.integers(2, 6))
    pts = rng.standard_normal((n, d)).astype(np.float32)
    t = float(rng.uniform(0.8, 2.5))
    r = oracle.run(pts, t, 2)
    ref = brute.brute_diagrams(pts, oracle.threshold_sq_f32(t), 2)
    for k in range(3):
        assert _bits_equal(brute.multiset(r["diagrams"][k]), brute.multiset(ref[k])), f"H{k}"


def test_planted_mistake_is_caught():
    """Dropping one cofacet from a column (a plausible indexing slip) changes
    the brute-force diagram comparison: the check has teeth."""
    pts = cases.randn(14, 3, 11)
    t = 1.7
    ref = brute.brute_diagrams(pts, oracle.threshold_sq_f32(t), 1)
    D = brute.all_dsq(pts)
    bad = D.copy()
    i, j = np.unravel_index(np.argmin(D + np.eye(14) * 1e9), D.shape)
    bad[i, j] = bad[j, i] = np.float32(0.5 * D[i, j])  # wrong term for one pair
    simp_ok = brute.brute_diagrams(pts, oracle.threshold_sq_f32(t), 1)
    assert all(_bits_equal(a, b) for a, b in zip(simp_ok, ref))
    orig = brute.all_dsq
    try:
        brute.all_dsq = lambda p: bad
        planted = brute.brute_diagrams(pts, oracle.threshold_sq_f32(t), 1)
    finally:
        brute.all_dsq = orig
    r = oracle.run(pts, t, 1)
    assert not all(_bits_equal(brute.multiset(r["diagrams"][k]), brute.multiset(planted[k]))
                   for k in range(2))


# ------------------------------------------------ reference known answers

def test_kat_circle_betti():
    # testing/test_persistence.py:476-492: circle n=20, t=0.5 -> Betti [1, 1]
    r = oracle.run(cases.circle(20), 0.5, 1)
    h0, h1 = r["diagrams"]
    assert int(np.isinf(h0[:, 1]).sum()) == 1
    assert int(np.isinf(h1[:, 1]).sum()) == 1


def test_kat_clifford_torus():
    # testing/test_persistence.py:541-570: n=200, t=1.4, max_dim=2 ->
    # 1 essential H0, 2 long H1 bars, 1 long H2 bar (persistence > 0.4)
    pts = cases.clifford(200, 42)
    r = oracle.run(pts, 1.4, 2)
    h0, h1, h2 = r["diagrams"]
    assert int(np.isinf(h0[:, 1]).sum()) == 1
    long1 = h1[(h1[:, 1] - h1[:, 0]) > 0.4]
    long2 = h2[(h2[:, 1] - h2[:, 0]) > 0.4]
    assert len(long1) == 2
    assert len(long2) == 1


def test_kat_sphere_h2():
    # testing/test_persistence.py:190-213: noisy 2-sphere, prominent H2 bar
    # (40 points: the reference's 20-point CUDA draw is not reproducible, and
    # 20 numpy-drawn points leave gaps that close only near the void's death)
    pts = cases.sphere(40, 3, 42, 0.05)
    r = oracle.run(pts, 2.0, 2)
    h2 = r["diagrams"][2]
    fin = h2[np.isfinite(h2[:, 1])]
    assert fin.shape[0] > 0 and float((fin[:, 1] - fin[:, 0]).max()) > 0.1


def test_kat_tiny_threshold_empty_h1():
    # testing/test_reduce_h1.py:177-183
    r = oracle.run(cases.randn(50, 3, 7), 1e-10, 1)
    assert r["diagrams"][1].shape == (0, 2)
    assert r["diagrams"][0].shape == (50, 2)


def test_kat_tight_cluster():
    # testing/test_reduce_h1.py:186-199: one component, no H1
    r = oracle.run(cases.randn(10, 2, 42, 0.01), 0.5, 1)
    assert int(np.isinf(r["diagrams"][0][:, 1]).sum()) == 1
    assert r["diagrams"][1].shape == (0, 2)


def test_row_order_rule():
    """H1 rows are sorted by birth, and rows of equal birth follow the
    reference's concatenation order (apparent asc, reduced desc, essential
    desc) -- checked on the grid with many ties."""
    r = oracle.run(cases.grid3(), 2.0, 1)
    h1 = r["diagrams"][1]
    assert np.all(np.diff(h1[:, 0]) >= 0)
    ess = h1[np.isinf(h1[:, 1])]
    fin = h1[np.isfinite(h1[:, 1])]
    assert ess.shape[0] + fin.shape[0] == h1.shape[0]


# ---------------------------------------------------------------- kNN threshold

def test_knn_oracle_vs_exact_brute():
    """k-th NN squared distance against exact-rational FMA chains and a full
    sort (no insertion list), on tiny clouds with ties."""
    for seed, (n, d) in enumerate([(9, 2), (12, 3), (7, 5)]):
        pts = cases.randn(n, d, 100 + seed)
        pts[3] = pts[1]  # a duplicate point: distance 0 counts
        D = brute.all_dsq(pts)
        for k in range(1, n):
            got = oracle.knn_kth_dsq(pts, k)
            for q in range(n):
                ds = sorted(D[q, j] for j in range(n) if j != q)
                assert got[q] == ds[k - 1], (seed, k, q)


def test_knn_oracle_vs_torch_cdist():
    """The reference's cdist + topk k-th distance agrees to cdist rounding."""
    import torch
    x = cases.randn(400, 6, 5)
    for k in (1, 20):
        ours = oracle.sqrt_f32(oracle.knn_kth_dsq(x, k))
        t = torch.from_numpy(x).double()
        dm = torch.cdist(t, t)
        dm.fill_diagonal_(float("inf"))
        ref = dm.topk(k, dim=1, largest=False).values[:, -1].numpy()
        assert np.allclose(ours, ref, rtol=1e-6, atol=0)


def test_quantile_matches_torch():
    """quantile_f32 is torch.quantile (the call the reference makes,
    threshold.py:77) on FP32 input, bit for bit."""
    import torch
    rng = np.random.default_rng(0)
    for n in (1, 2, 3, 10, 999, 1000, 4097):
        v = rng.gamma(2.0, 1.0, n).astype(np.float32)
        for q in (0.0, 0.05, 0.25, 0.5, 0.9, 0.95, 0.999, 1.0, 1 / 3):
            ref = torch.quantile(torch.from_numpy(v), q).item()
            assert oracle.quantile_f32(v, q) == np.float32(ref), (n, q)


def test_auto_threshold_oracle_edge_cases():
    assert oracle.auto_threshold(np.zeros((0, 3), np.float32)) == 0.0
    assert oracle.auto_threshold(np.zeros((1, 3), np.float32)) == 0.0
    # k clamps to n-1; percentile clamps to [0, 100]
    x = cases.randn(5, 2, 1)
    assert oracle.auto_threshold(x, k=50) == oracle.auto_threshold(x, k=4)
    assert oracle.auto_threshold(x, percentile=500) == oracle.auto_threshold(x, percentile=100)
    # a regular polygon: every point's k-th neighbour is at the same distance
    th = np.arange(12) * (2 * np.pi / 12)
    poly = np.column_stack([np.cos(th), np.sin(th)]).astype(np.float32)
    t = oracle.auto_threshold(poly, k=2, percentile=50)
    assert abs(t - 2 * np.sin(np.pi / 12)) < 1e-6


def test_enclosing_radius_oracle():
    """Against exact-rational distances, and against float64 cdist."""
    import torch
    x = cases.randn(11, 3, 4)
    D = brute.all_dsq(x)
    np.fill_diagonal(D, 0.0)
    ref = float(np.sqrt(np.float64(D.max(axis=1).min())).astype(np.float32))
    assert oracle.enclosing_radius(x) == ref
    y = cases.randn(300, 7, 5)
    t = torch.from_numpy(y).double()
    r64 = torch.cdist(t, t).max(dim=1).values.min().item()
    assert abs(oracle.enclosing_radius(y) - r64) <= 1e-6 * r64
    assert oracle.enclosing_radius(np.zeros((1, 2), np.float32)) == 0.0


# ---------------------------------------------- reduction store semantics

@pytest.mark.parametrize("seed", range(6))
def test_exact_mode_matches_brute_homology(seed):
    """The complete reduction (exact=True) is pinned by brute-force homology
    like the reference mode."""
    rng = np.random.default_rng(2000 + seed)
    n = int(rng.integers(8, 16))
    pts = rng.standard_normal((n, 3)).astype(np.float32)
    t = float(rng.uniform(1.0, 2.5))
    r = oracle.run(pts, t, 1, exact=True)
    ref = brute.brute_diagrams(pts, oracle.threshold_sq_f32(t), 1)
    for k in range(2):
        assert _bits_equal(brute.multiset(r["diagrams"][k]), brute.multiset(ref[k])), f"H{k}"


def test_store_overflow_changes_only_reference_mode():
    """A cloud of the uneven-batch workload (n=1184, d=10, auto threshold)
    overflows the reference's store (reduce_h1_gpu.py:286-289): the
    reference mode drops columns, the exact mode does not, and where
    nothing is dropped the two modes agree bit for bit."""
    x = cases.batch_uneven(256, seed=0)[140]
    t = oracle.auto_threshold(x)
    ref = oracle.run(x, t, 1)
    ex = oracle.run(x, t, 1, exact=True)
    assert ref["stats"]["pool_drops"] > 0 and ex["stats"]["pool_drops"] == 0
    assert ref["stats"]["residual1"] == ex["stats"]["residual1"]
    # H0 and the apparent pairs do not depend on the store
    assert np.array_equal(ref["h0"], ex["h0"])
    assert np.array_equal(ref["apparent1"], ex["apparent1"])
    y = cases.randn(200, 4, 5)
    a, b = oracle.run(y, 1.0, 1), oracle.run(y, 1.0, 1, exact=True)
    assert a["stats"]["pool_drops"] == 0
    assert np.array_equal(a["h1"].view(np.uint32), b["h1"].view(np.uint32))


def test_max_dim2_h1_is_complete_reduction():
    """For max_dim=2 the reference's H1 comes from giotto-ph (a complete
    reduction): the oracle's H1 then equals its exact-mode H1."""
    x = cases.randn(120, 3, 9)
    a = oracle.run(x, 1.1, 2)
    b = oracle.run(x, 1.1, 1, exact=True)
    assert np.array_equal(a["h1"].view(np.uint32), b["h1"].view(np.uint32))


@pytest.mark.parametrize("max_col,max_pool", [(4, 0), (1 << 20, 60), (6, 80), (10, 300)])
@pytest.mark.parametrize("seed", range(3))
def test_reference_limits_match_python_restatement(seed, max_col, max_pool):
    """The reference's MAX_COL truncation and store cap, lowered so that tiny
    clouds reach them: the oracle's pairs and essentials equal a second,
    independent restatement of the reference's loop (tests/brute.py)."""
    rng = np.random.default_rng(seed)
    pts = rng.standard_normal((60, 3)).astype(np.float32)
    t = 1.3
    o = oracle.run(pts, t, 1, max_col=max_col, max_pool=max_pool)
    E, T = o["stats"]["E"], o["stats"]["T"]
    mst = np.zeros(E, dtype=bool)
    mst[o["mst_idx"]] = True
    ef = oracle.sqrt_f32(o["edge_dsq"])
    r2f = np.empty(E + T, dtype=np.float32)
    r2f[:E] = ef
    r2f[o["tri_global_rank"]] = ef[o["tri_max_rank"]]
    pairs, ess, drops, truncs = brute.reference_reduction(
        o["cof_off"], o["cof_data"], mst, r2f, max_col, max_pool or 100_000_000)
    assert list(zip(o["h1_pb"].tolist(), o["h1_pd"].tolist())) == pairs
    assert o["h1_eb"].tolist() == ess
    if max_col < 1 << 20:
        assert truncs > 0 and o["stats"]["truncations"] > 0
    if max_pool:
        assert drops > 0 and o["stats"]["pool_drops"] > 0


@pytest.mark.parametrize("n,d,t,seed", [(700, 3, 0.5, 0), (300, 1, 0.05, 1), (2500, 2, 0.06, 2)])
def test_edge_list_entry_matches_enumeration(n, d, t, seed):
    """orc_run_ex from a shuffled edge list rebuilds the enumeration's
    emission order: every output equals the full run's."""
    rng = np.random.default_rng(seed)
    pts = rng.standard_normal((n, d)).astype(np.float32)
    if d == 1:
        pts = np.round(pts * 50) / 50   # many equal distances
    a = oracle.run(pts, t, 2)
    p = rng.permutation(a["stats"]["E"])
    b = oracle.run(pts, t, 2, edges=(a["edge_i"][p], a["edge_j"][p], a["edge_dsq"][p]))
    for k in oracle.NAMES:
        if k == "stats":
            continue
        assert np.array_equal(a[k].view(np.uint8), b[k].view(np.uint8)), k


def test_wrapped_key_order_on_grid():
    """n = 47^3 = 103,823 > 92,682: the packed colex field wraps, equal dsq
    values are ordered by the wrapped colex, and fully tied keys by the
    emission order.  The oracle's edge order equals numpy's restatement."""
    pts = cases.grid_points(47)
    i, j, dsq = cases.grid_edges(pts, 1.5)
    assert cases.wrapped_key_ties(i, j, dsq) > 0
    assert (j.astype(np.int64) * (j - 1) // 2 + i >= 2**32).any()
    p = np.random.default_rng(0).permutation(i.size)
    o = oracle.run(pts, 1.5, 0, edges=(i[p], j[p], dsq[p]))
    order = cases.reference_edge_order(i, j, dsq)
    assert np.array_equal(o["edge_i"], i[order])
    assert np.array_equal(o["edge_j"], j[order])
    assert np.array_equal(o["edge_dsq"].view(np.uint32), dsq[order].view(np.uint32))
