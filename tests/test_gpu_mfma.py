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
