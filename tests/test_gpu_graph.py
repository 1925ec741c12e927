"""GPU: the hipGraph-captured H0 pipeline equals the oracle and the eager
path on every replay, with eager work between replays, across a capacity
re-capture, and through the split-bf16 screen."""
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
    got, want = np.asarray(got), np.asarray(want)
    assert got.shape == want.shape, f"{name}: shape {got.shape} vs {want.shape}"
    assert np.array_equal(_bits(got), _bits(want)), f"{name}: values differ"


def _check(g, pts, t):
    import oracle
    from flash_ph import rips_persistence
    ref = oracle.run(pts, t, 0)
    h0 = g(torch.from_numpy(pts).cuda())
    # eager work between replays, and the eager path itself
    eager = rips_persistence(torch.from_numpy(pts).cuda(), t, max_dim=0)[0]
    torch.cuda.synchronize()
    _eq("h0", h0.cpu().numpy(), ref["h0"])
    _eq("h0_eager", eager.cpu().numpy(), ref["h0"])
    ei, ej, dsq = g.edges()
    _eq("edge_i", ei.cpu().numpy(), ref["edge_i"])
    _eq("edge_j", ej.cpu().numpy(), ref["edge_j"])
    _eq("edge_dsq", dsq.cpu().numpy(), ref["edge_dsq"])
    ptr = g.adj_ptr.cpu().numpy()
    _eq("adj_ptr", ptr, ref["adj_ptr"])
    _eq("adj_idx", g.adj_idx[:ptr[-1]].cpu().numpy(), ref["adj_idx"])
    _eq("adj_rank", g.adj_rank[:ptr[-1]].cpu().numpy(), ref["adj_rank"])


def test_replays_match_oracle():
    from flash_ph.graph import RipsH0Graph
    t = 2.2
    g = RipsH0Graph(2000, 10, t)
    for seed in (1, 2, 3, 1, 4):
        print(f"[gpu] graph replay seed {seed}", flush=True)
        _check(g, cases.randn(2000, 10, seed), t)
    assert g.captures == 1
    g.close()


def test_recapture_on_overflow():
    from flash_ph.graph import RipsH0Graph
    t = 0.6
    g = RipsH0Graph(1000, 4, t, edge_capacity=64)
    pts = cases.clifford(1000, 5, 0.02)
    _check(g, pts, t)
    assert g.captures == 2 and g.cap >= g.E > 64
    _check(g, cases.clifford(1000, 6, 0.02), t)
    g.close()


def test_split_bf16_screen_graph():
    from flash_ph.graph import RipsH0Graph
    t = 6.3
    g = RipsH0Graph(4096, 32, t)
    assert g.screen == 1
    for seed in (11, 12):
        _check(g, cases.randn(4096, 32, seed), t)
    g.close()


def test_no_edges_and_shape_errors():
    from flash_ph.graph import RipsH0Graph
    g = RipsH0Graph(50, 3, 1e-10)
    _check(g, cases.randn(50, 3, 7), 1e-10)
    with pytest.raises(ValueError):
        g(torch.zeros(49, 3))
    g.close()
