"""GPU parity at the large configurations and on the reference's limit
branches:

* n = 47^3 = 103,823 > 92,682 integer-grid points: the wrapped colex field of
  the sort key and the emission-order tie break (rips.py:30-50,188-202,
  kernels/edge_kernel.py:133-149,199-210) against numpy's restatement;
* n = 200,000, d = 50, N(0,1), t = 6.0, H0+H1 (the split-bf16 workload):
  split-bf16 and direct screens give identical edges; sampled vertex rows
  equal the oracle's FMA-chain neighbours; edge order, forest, H0 and H1
  equal the oracle run from the same edge set;
* n = 1,000, d = 10, N(0,1), auto_threshold(k=20, p=95), H0+H1: the dense
  CPU-baseline workload, bit-exact against the oracle;
* the reference's MAX_COL truncation and store cap, lowered so that small
  clouds reach them, against the oracle with the same limits.
"""
from __future__ import annotations

import numpy as np
import pytest
import torch

import cases

pytestmark = pytest.mark.gpu


def _bits(a):
    a = np.ascontiguousarray(a)
    return a.view(np.uint32) if a.dtype == np.float32 else a


def _eq(name, got, want):
    got = np.asarray(got)
    want = np.asarray(want)
    assert got.shape == want.shape, f"{name}: shape {got.shape} vs {want.shape}"
    assert np.array_equal(_bits(got), _bits(want)), f"{name}: values differ"


def _cpu(t):
    return t.detach().cpu().numpy()


def test_wrapped_key_order_grid():
    import oracle
    from flash_ph import _pipeline, rips_persistence
    pts = cases.grid_points(47)
    t = 1.5
    print(f"[gpu] grid n={pts.shape[0]} t={t}", flush=True)
    i, j, dsq = cases.grid_edges(pts, t)
    assert cases.wrapped_key_ties(i, j, dsq) > 0
    x = torch.from_numpy(pts).cuda()
    cx = _pipeline.build_edges(x, t)
    torch.cuda.synchronize()
    order = cases.reference_edge_order(i, j, dsq)
    _eq("edge_i", _cpu(cx.edge_i), i[order])
    _eq("edge_j", _cpu(cx.edge_j), j[order])
    _eq("edge_dsq", _cpu(cx.edge_dsq), dsq[order])
    ref = oracle.run(pts, t, 0, edges=(i, j, dsq))
    _eq("adj_idx", _cpu(cx.adj_idx), ref["adj_idx"])
    _eq("adj_rank", _cpu(cx.adj_rank), ref["adj_rank"])
    h0 = rips_persistence(x, t, max_dim=0)[0]
    _eq("h0", _cpu(h0), ref["h0"])


CFG4 = dict(n=200_000, d=50, t=6.0, seed=0)


@pytest.fixture(scope="module")
def cfg4_points():
    rng = np.random.default_rng(CFG4["seed"])
    return rng.standard_normal((CFG4["n"], CFG4["d"])).astype(np.float32)


def test_cfg4_screens_agree(cfg4_points):
    """Split-bf16 matrix-core screen vs the direct FP32 screen at full size:
    same edges in the same order, same dsq bits."""
    from flash_ph import _pipeline
    x = torch.from_numpy(cfg4_points).cuda()
    old = _pipeline.SCREEN
    try:
        print("[gpu] n=200000 d=50 mfma screen", flush=True)
        _pipeline.SCREEN = "mfma"
        a = _pipeline.build_edges(x, CFG4["t"])
        print("[gpu] n=200000 d=50 direct screen", flush=True)
        _pipeline.SCREEN = "direct"
        b = _pipeline.build_edges(x, CFG4["t"])
    finally:
        _pipeline.SCREEN = old
    torch.cuda.synchronize()
    assert a.E == b.E and a.E > 100_000
    for k in ("edge_i", "edge_j", "edge_dsq", "adj_ptr", "adj_idx", "adj_rank"):
        _eq(k, _cpu(getattr(a, k)), _cpu(getattr(b, k)))
    assert int(a.info["screen_candidates"].item()) >= a.E


def test_cfg4_h0_h1_vs_oracle(cfg4_points):
    """H0+H1 of the n=200,000, d=50 cloud in the reference's semantics: the
    oracle starts from the GPU's edge set (its enumeration would take hours
    on one core); the set itself is checked row by row on sampled vertices
    against the oracle's FMA-chain distances."""
    import oracle
    from flash_ph import _pipeline
    pts = cfg4_points
    t = CFG4["t"]
    x = torch.from_numpy(pts).cuda()
    print("[gpu] n=200000 d=50 H0+H1", flush=True)
    dg, cx = _pipeline.compute(x, t, 1, semantics="reference")
    torch.cuda.synchronize()
    ei, ej, ed = _cpu(cx.edge_i), _cpu(cx.edge_j), _cpu(cx.edge_dsq)
    ptr, idx = _cpu(cx.adj_ptr), _cpu(cx.adj_idx)
    adsq = _cpu(cx.adj_dsq)
    rng = np.random.default_rng(7)
    deg = np.diff(ptr)
    sample = np.concatenate([rng.choice(pts.shape[0], 24, replace=False),
                             np.argsort(deg)[-8:]])   # include the busiest rows
    for v in sample:
        oi, od = oracle.vertex_neighbours(pts, t, int(v))
        _eq(f"row {v} idx", idx[ptr[v]:ptr[v + 1]], oi)
        _eq(f"row {v} dsq", adsq[ptr[v]:ptr[v + 1]], od)
    print("[gpu] oracle from the edge set", flush=True)
    perm = rng.permutation(ei.size)
    ref = oracle.run(pts, t, 1, edges=(ei[perm], ej[perm], ed[perm]))
    _eq("edge_i", ei, ref["edge_i"])
    _eq("edge_j", ej, ref["edge_j"])
    _eq("edge_dsq", ed, ref["edge_dsq"])
    assert (ej.astype(np.int64) * (ej - 1) // 2 + ei >= 2**32).any()  # wrapped keys occur
    mst = np.nonzero(_cpu(cx.mst_mask)[:cx.E])[0]
    _eq("mst_idx", mst, ref["mst_idx"])
    _eq("h0", _cpu(dg[0]), ref["h0"])
    _eq("h1", _cpu(dg[1]), ref["h1"])


def test_cfg1_dense_auto_threshold():
    """n=1,000, d=10, N(0,1) (default_rng(0)), t = auto_threshold(k=20,
    p=95) ~ 3.37: E ~ 79K, T ~ 2.76M; H0+H1 bit-exact against the oracle in
    the reference's semantics (threshold.py:20-78, reduce_h1_gpu.py:438-604)."""
    import oracle
    from flash_ph import auto_threshold, rips_persistence
    pts = cases.randn(1000, 10, 0)
    x = torch.from_numpy(pts).cuda()
    t = auto_threshold(x, k=20, percentile=95.0)
    t_ref = oracle.auto_threshold(pts, 20, 95.0)
    assert np.float32(t).view(np.uint32) == np.float32(t_ref).view(np.uint32)
    print(f"[gpu] n=1000 d=10 t={t} H0+H1", flush=True)
    dg = rips_persistence(x, t, max_dim=1)
    ref = oracle.run(pts, t, 1)
    assert ref["stats"]["T"] > 2_000_000
    _eq("h0", _cpu(dg[0]), ref["h0"])
    _eq("h1", _cpu(dg[1]), ref["h1"])


# (name, points, t, max_col, max_pool): lowered MAX_COL (truncation), a
# lowered store cap (dropped columns), both; the first two clouds have few
# enough rows for the replay check, which must then hand over to the serial
# reducer; the last one's rows exceed it.
LIMIT_CASES = [
    ("trunc", cases.randn(60, 3, 0), 1.3, 4, 0),
    ("drops", cases.randn(60, 3, 1), 1.3, 1 << 20, 60),
    ("both", cases.randn(300, 4, 2), 1.2, 12, 2000),
    ("trunc_big", cases.randn(700, 3, 3), 0.9, 24, 0),
]


@pytest.mark.parametrize("name,pts,t,max_col,max_pool", LIMIT_CASES,
                         ids=[c[0] for c in LIMIT_CASES])
def test_reference_limits(name, pts, t, max_col, max_pool):
    import oracle
    from flash_ph import _pipeline, rips_persistence
    old = _pipeline.REF_MAX_COL, _pipeline.REF_MAX_POOL
    try:
        _pipeline.REF_MAX_COL = max_col
        if max_pool:
            _pipeline.REF_MAX_POOL = max_pool
        print(f"[gpu] limits {name} n={pts.shape[0]} max_col={max_col} max_pool={max_pool}",
              flush=True)
        ref = oracle.run(pts, t, 1, max_col=max_col, max_pool=max_pool)
        if max_col < 1 << 20:
            assert ref["stats"]["truncations"] > 0
        if max_pool:
            assert ref["stats"]["pool_drops"] > 0
        dg = rips_persistence(torch.from_numpy(pts).cuda(), t, max_dim=1)
        _eq("h1", _cpu(dg[1]), ref["h1"])
    finally:
        _pipeline.REF_MAX_COL, _pipeline.REF_MAX_POOL = old
