"""The host (CPU) pipeline, libflashph_cpu, against the oracle on the CPU:
diagrams bit for bit on the parity families (both reduction semantics,
max_dim 0..2), the reference's lowered limits, the filtration and its CSR,
and thread-count independence."""
from __future__ import annotations

import numpy as np
import pytest
import torch

import cases
import oracle
from flash_ph import _cpu


def _eq(a, b):
    return a.shape == b.shape and np.array_equal(np.ascontiguousarray(a).view(np.uint32),
                                                 np.ascontiguousarray(b).view(np.uint32))


@pytest.mark.parametrize("name,make,t,md", cases.SMALL + cases.MEDIUM[:4],
                         ids=[c[0] for c in cases.SMALL + cases.MEDIUM[:4]])
@pytest.mark.parametrize("exact", [False, True])
def test_cpu_pipeline_matches_oracle(name, make, t, md, exact):
    pts = make()
    thr = oracle.threshold_sq_f32(t)
    for max_dim in sorted({0, 1, md}):
        got = _cpu.rips(pts, thr, max_dim, exact_h1=exact)
        ref = oracle.run(pts, t, max_dim, exact=exact)["diagrams"]
        for k in range(max_dim + 1):
            assert _eq(got[k], ref[k]), (name, max_dim, k)


@pytest.mark.parametrize("max_col,max_pool", [(4, 0), (1 << 20, 60), (12, 2000)])
def test_cpu_reference_limits(max_col, max_pool):
    pts = cases.randn(300, 4, 2)
    got = _cpu.rips(pts, oracle.threshold_sq_f32(1.2), 1, max_col=max_col, max_pool=max_pool)
    ref = oracle.run(pts, 1.2, 1, max_col=max_col, max_pool=max_pool)
    assert _eq(got[1], ref["h1"])


def test_cpu_filtration_and_api():
    from flash_ph.rips import rips_filtration
    pts = cases.randn(400, 3, 8)
    f = rips_filtration(torch.from_numpy(pts), 0.8)
    ref = oracle.run(pts, 0.8, 0)
    for k, r in [("edge_i", "edge_i"), ("edge_j", "edge_j"), ("edge_dist_sq", "edge_dsq"),
                 ("vert_adj_ptr", "adj_ptr"), ("vert_adj_idx", "adj_idx"),
                 ("vert_adj_rank", "adj_rank"), ("vert_adj_dist_sq", "adj_dsq")]:
        assert _eq(getattr(f, k).numpy(), ref[r]), k
    assert f.edge_i.device.type == "cpu"


def test_cpu_threads_do_not_change_results():
    pts = cases.randn(2000, 10, 1)
    thr = oracle.threshold_sq_f32(2.2)
    a = _cpu.rips(pts, thr, 2, threads=1)
    b = _cpu.rips(pts, thr, 2, threads=8)
    for x, y in zip(a, b):
        assert _eq(x, y)


def test_cpu_wrapped_key_order():
    """n > 92,682: the emission-order tie break of the wrapped key, and the
    exact colex order, on the 47^3 grid."""
    pts = cases.grid_points(47)
    i, j, dsq = cases.grid_edges(pts, 1.5)
    f = _cpu.filtration(pts, oracle.threshold_sq_f32(1.5))
    order = cases.reference_edge_order(i, j, dsq)
    assert np.array_equal(f["edge_i"], i[order]) and np.array_equal(f["edge_j"], j[order])
    g = _cpu.filtration(pts, oracle.threshold_sq_f32(1.5), exact_colex=True)
    o2 = np.lexsort((i, j, dsq.view(np.uint32)))
    assert np.array_equal(g["edge_i"], i[o2]) and np.array_equal(g["edge_j"], j[o2])


SCAN_CASES = [("randn_d3_n300", cases.randn(300, 3, 1), 0.7),
              ("clifford_300", cases.clifford(300, 2, 0.02), 0.6),
              ("sphere_d4_n250", cases.sphere(250, 4, 3, 0.05), 0.7),
              ("dup_points", np.repeat(cases.randn(60, 2, 3), 2, axis=0), 0.5)]


@pytest.mark.parametrize("name,pts,t", SCAN_CASES, ids=[c[0] for c in SCAN_CASES])
def test_cpu_apparent_scans_match_definitions(name, pts, t):
    """Host scans (the reference's Numba paths) against the predicates
    written from their definitions; for H2 also the reference's d < s case
    rule, which must agree with "largest face of the youngest cofacet"."""
    import scan_defs
    from flash_ph.kernels.apparent_pair_h2_kernel import h2_apparent_pair_scan
    from flash_ph.kernels.apparent_pair_kernel import apparent_pair_scan
    f = _cpu.filtration(pts, oracle.threshold_sq_f32(t))
    ei, ej = f["edge_i"], f["edge_j"]
    ptr, idx, rank = f["adj_ptr"], f["adj_idx"], f["adj_rank"]
    ranks = np.arange(ei.size, dtype=np.int32)
    tt = [torch.from_numpy(np.ascontiguousarray(x)) for x in (ei, ej, ranks, ptr, idx, rank)]
    flag, colex = apparent_pair_scan(*tt, 0)
    wf, wc = scan_defs.h1_definition(ei, ej, ranks, ptr, idx, rank)
    assert np.array_equal(flag.numpy(), wf) and np.array_equal(colex.numpy(), wc)
    assert wf.sum() > 0
    tri = scan_defs.triangles(ei, ej, ptr, idx, rank)
    cols = [torch.from_numpy(np.ascontiguousarray(tri[:, q]).astype(np.int32)) for q in range(4)]
    f2, c2 = h2_apparent_pair_scan(*cols, *tt[3:], 0)
    wf2, wc2 = scan_defs.h2_definition(tri, ptr, idx, rank)
    assert np.array_equal(f2.numpy(), wf2) and np.array_equal(c2.numpy(), wc2)
    assert 0 < wf2.sum() < len(tri)
    # the reference's case rule on the unbounded walk gives the same set
    big = 1 << 40
    for k in range(len(tri)):
        a, b, c, rt = (int(x) for x in tri[k])
        d = scan_defs.h2_reference_walk(a, b, c, rt, ptr, idx, rank, big)
        app = d >= 0 and scan_defs.h2_oldest_face(a, b, c, rt, d, ptr, idx, rank)
        assert app == bool(wf2[k]), (name, k)


@pytest.mark.parametrize("idx", [140, 225])
def test_cpu_store_overflow_default_capacity(idx):
    """Clouds of the uneven-batch workload whose reference store overflows at
    its own capacity (sized from 3 cofacet entries per triangle, forest
    columns included): the host pipeline drops the oracle's columns, makes
    its additions and gives its diagram."""
    x = cases.batch_uneven(256, seed=0)[idx]
    t = oracle.auto_threshold(x)
    ref = oracle.run(x, t, 1)
    assert ref["stats"]["pool_drops"] > 0
    dg, st = _cpu.rips(x, oracle.threshold_sq_f32(t), 1, want_stats=True)
    assert st["drops"] == ref["stats"]["pool_drops"] and st["xor1"] == ref["stats"]["xor1"]
    assert _eq(dg[0], ref["h0"]) and _eq(dg[1], ref["h1"])
