"""The split-bf16 matrix-core screen gives exactly the direct screen's edges
(same pairs, same dsq bits), for dimensions around the MFMA k-step (16) and
cloud sizes around the 128-point tile, and the whole pipeline's diagrams are
unchanged.  Also reports the FP64 error of the FP32 distances."""
from __future__ import annotations

import numpy as np
import pytest
import torch

import cases

pytestmark = pytest.mark.gpu


def _edges(pts, t, mode):
    from flash_ph import _pipeline
    old = _pipeline.SCREEN
    _pipeline.SCREEN = mode
    try:
        cx = _pipeline.build_edges(pts, t)
    finally:
        _pipeline.SCREEN = old
    return cx


@pytest.mark.parametrize("n,d,t", [(300, 32, 7.0), (1000, 50, 8.6), (777, 64, 10.0),
                                   (129, 70, 11.0), (4100, 50, 8.4), (2000, 33, 7.2),
                                   (513, 16, 4.9), (600, 3, 0.6)])
def test_mfma_screen_equals_direct(n, d, t):
    print(f"[gpu] mfma screen n={n} d={d}", flush=True)
    x = torch.from_numpy(cases.randn(n, d, 11 + d)).cuda()
    a = _edges(x, t, "direct")
    b = _edges(x, t, "mfma")
    assert a.E == b.E and a.E > 0
    for name in ("edge_i", "edge_j", "edge_dsq"):
        assert torch.equal(getattr(a, name)[:a.E].view(torch.int32),
                           getattr(b, name)[:b.E].view(torch.int32)), name
    cand = int(b.info["screen_candidates"].item())
    assert cand >= b.E


def test_mfma_pipeline_diagrams_and_fp64_error():
    from flash_ph import _pipeline, rips_persistence
    print("[gpu] mfma pipeline", flush=True)
    x0 = cases.randn(6000, 50, 5)
    x = torch.from_numpy(x0).cuda()
    _pipeline.SCREEN = "direct"
    ref = rips_persistence(x, 8.0, max_dim=1)
    _pipeline.SCREEN = "auto"
    got = rips_persistence(x, 8.0, max_dim=1)
    for a, b in zip(ref, got):
        assert torch.equal(a.view(torch.int32), b.view(torch.int32))
    cx = _edges(x, 8.0, "auto")
    assert cx.info["screen"] == "mfma"
    # FP64 error of the FP32 direct-form distances (the filtration values)
    i, j = cx.edge_i[:cx.E].long(), cx.edge_j[:cx.E].long()
    x64 = x.double()
    d64 = ((x64[i] - x64[j]) ** 2).sum(1)
    rel = ((cx.edge_dsq[:cx.E].double() - d64).abs() / d64).max().item()
    assert rel < 50 * 2.0 ** -24


@pytest.mark.parametrize("d", [50, 64, 70])
def test_mfma_near_threshold_pairs(d):
    """Pairs whose FP32 dsq lies within +-5 ulp of t^2 (both sides): the
    split-bf16 screen keeps exactly the direct screen's and the oracle's
    edges -- the error margin admits every pair the direct form keeps."""
    import oracle
    pts, offs = cases.near_threshold_pairs(300, d, 3.0, seed=d)
    assert (offs < 0).sum() > 50 and (offs >= 0).sum() > 50
    print(f"[gpu] near-threshold pairs d={d}", flush=True)
    x = torch.from_numpy(pts).cuda()
    a = _edges(x, 3.0, "direct")
    b = _edges(x, 3.0, "mfma")
    ref = oracle.run(pts, 3.0, 0)
    assert a.E == b.E == ref["stats"]["E"]
    for name in ("edge_i", "edge_j", "edge_dsq"):
        got = getattr(b, name)[:b.E].cpu().numpy()
        assert np.array_equal(got.view(np.int32), ref[name].view(np.int32)), name
        assert torch.equal(getattr(a, name)[:a.E].view(torch.int32),
                           getattr(b, name)[:b.E].view(torch.int32)), name


def test_mfma_offset_cloud():
    """A cloud far from the origin (N(0,1) + 1000 in every coordinate): the
    screen works on centred coordinates, so the margin band stays small and
    the edges equal the direct screen's."""
    print("[gpu] mfma offset cloud", flush=True)
    pts = (cases.randn(6000, 48, 21) + np.float32(1000.0)).astype(np.float32)
    x = torch.from_numpy(pts).cuda()
    a = _edges(x, 7.5, "direct")
    b = _edges(x, 7.5, "mfma")
    assert a.E == b.E and a.E > 1000
    for name in ("edge_i", "edge_j", "edge_dsq"):
        assert torch.equal(getattr(a, name)[:a.E].view(torch.int32),
                           getattr(b, name)[:b.E].view(torch.int32)), name
    cand = int(b.info["screen_candidates"].item())
    assert b.E <= cand <= 2 * b.E + 1000, (cand, b.E)


def test_fp64_validation_mode():
    """Validation mode: the band listed by the split-bf16 screen is exactly
    candidates minus edges; in FP64 every edge is an FP32 member and no band
    pair is; on the near-threshold cloud some pairs change membership in
    exact arithmetic (what the report is for), within a few ulp."""
    from flash_ph import _pipeline
    pts, offs = cases.near_threshold_pairs(300, 50, 3.0, seed=50)
    x = torch.from_numpy(pts).cuda()
    old = _pipeline.SCREEN
    try:
        _pipeline.SCREEN = "mfma"
        print("[gpu] fp64 validation near-threshold", flush=True)
        rep = _pipeline.fp64_screen_report(x, 3.0)
        cx = _edges(x, 3.0, "mfma")
    finally:
        _pipeline.SCREEN = old
    assert rep["screen"] == "mfma" and rep["edges"] == cx.E
    assert rep["edges"] + rep["band"] == int(cx.info["screen_candidates"].item())
    assert rep["edges_fp32_members"] == rep["edges"] and rep["band_fp32_members"] == 0
    assert rep["band"] > 0
    assert 0.0 < rep["max_ulp_fp32_vs_fp64"] < 64.0
    # the planted pairs within +-5 ulp of t^2: exact membership differs for some
    assert rep["fp32_in_exact_out"] + rep["exact_in_fp32_out"] >= 1
    print("[gpu] fp64 validation n=20000 d=40", flush=True)
    y = torch.from_numpy(cases.randn(20_000, 40, 3)).cuda()
    rep2 = _pipeline.fp64_screen_report(y, 5.2)
    assert rep2["screen"] == "mfma" and rep2["edges"] > 1000
    assert rep2["edges_fp32_members"] == rep2["edges"] and rep2["band_fp32_members"] == 0
