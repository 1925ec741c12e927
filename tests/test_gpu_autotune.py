"""GPU: every autotune candidate gives the oracle's output bit for bit, and
the tuner installs, reports and caches a valid choice."""
from __future__ import annotations

import numpy as np
import pytest
import torch

import cases

pytestmark = pytest.mark.gpu


@pytest.fixture(autouse=True)
def _clean_table():
    from flash_ph import autotune
    autotune.reset()
    yield
    autotune.reset()


def _bits(a):
    a = np.ascontiguousarray(a)
    return a.view(np.uint32) if a.dtype == np.float32 else a


def _eq(name, got, want):
    got, want = np.asarray(got), np.asarray(want)
    assert got.shape == want.shape, f"{name}: shape {got.shape} vs {want.shape}"
    assert np.array_equal(_bits(got), _bits(want)), f"{name}: values differ"


# ragged tile edges (n not a multiple of 32/64/128), d beyond one LDS chunk,
# a dense case and an n=10,000, d=10 one
SCREEN_CASES = [
    ("randn_n777_d3", lambda: cases.randn(777, 3, 5), 0.4),
    ("randn_n300_d50", lambda: cases.randn(300, 50, 6), 9.5),
    ("randn_n130_d10_dense", lambda: cases.randn(130, 10, 7), 6.0),
    ("randn_n5000_d10", lambda: cases.randn(5000, 10, 8), 2.0),
]


@pytest.mark.parametrize("tile", [32, 64, 128])
@pytest.mark.parametrize("name,make,t", SCREEN_CASES, ids=[c[0] for c in SCREEN_CASES])
def test_screen_tiles_match_oracle(tile, name, make, t):
    import oracle
    from flash_ph import _pipeline, autotune
    pts = make()
    n, d = pts.shape
    autotune.set_value("screen_tile", n, d, tile)
    assert autotune.query("screen_tile", n, d) == tile
    ref = oracle.run(pts, t, 0)
    cx = _pipeline.build_edges(torch.from_numpy(pts).cuda(), t)
    torch.cuda.synchronize()
    assert cx.info["screen"] == "direct"
    _eq("edge_i", cx.edge_i.cpu().numpy(), ref["edge_i"])
    _eq("edge_j", cx.edge_j.cpu().numpy(), ref["edge_j"])
    _eq("edge_dsq", cx.edge_dsq.cpu().numpy(), ref["edge_dsq"])


@pytest.mark.parametrize("knob", ["enum_block", "apparent_block"])
@pytest.mark.parametrize("value", [64, 128, 512, 1024])
def test_block_sizes_match_oracle(knob, value):
    import oracle
    from flash_ph import _pipeline, autotune
    pts, t = cases.randn(2000, 10, 1), 2.2
    ref = oracle.run(pts, t, 2)
    x = torch.from_numpy(pts).cuda()
    _, cx0 = _pipeline.compute(x, t, 2, semantics="reference")
    for size in (cx0.E, cx0.T):
        autotune.set_value(knob, size, 0, value)
    dgms, cx = _pipeline.compute(x, t, 2, semantics="reference")
    torch.cuda.synchronize()
    assert autotune.query(knob, cx.E) == value
    for k, name in enumerate(("h0", "h1", "h2")):
        _eq(name, dgms[k].cpu().numpy(), ref[name])


QUEUE_CASES = [
    ("randn_d10_n2000", lambda: cases.randn(2000, 10, 1), 2.2, 2, "reference"),
    ("clifford_1000_noisy", lambda: cases.clifford(1000, 3, 0.02), 0.45, 2, "reference"),
    ("o3_1024_exact", lambda: cases.o3(1024, 42), 1.8, 1, "exact"),
    ("randn_d10_n10000_t16", lambda: cases.randn(10000, 10, 0), 1.6, 2, "reference"),
]


@pytest.mark.parametrize("wgs", [4, 16, 64])
@pytest.mark.parametrize("name,make,t,max_dim,sem", QUEUE_CASES, ids=[c[0] for c in QUEUE_CASES])
def test_queue_workgroups_match_oracle(wgs, name, make, t, max_dim, sem, monkeypatch):
    """The residual reduction's queue spread over several workgroups (device-
    coherent publication, pivots rebuilt from the owner table) gives the
    oracle's diagrams."""
    import oracle
    from flash_ph import _pipeline, autotune
    monkeypatch.setattr(_pipeline, "_WAIT_SECONDS", 10)
    print(f"[gpu] queue_wgs={wgs} {name}", flush=True)
    pts = make()
    ref = oracle.run(pts, t, max_dim, exact=(sem == "exact"))
    x = torch.from_numpy(pts).cuda()
    _, cx0 = _pipeline.compute(x, t, max_dim, semantics=sem)
    for size in (cx0.E, cx0.T):
        autotune.set_value("queue_wgs", size, 0, wgs)
    for rep in range(2):
        dgms, cx = _pipeline.compute(x, t, max_dim, semantics=sem)
        torch.cuda.synchronize()
        for k in range(max_dim + 1):
            name_k = ("h0", "h1", "h2")[k]
            _eq(f"{name_k} rep {rep}", dgms[k].cpu().numpy(), ref[name_k])
        if max_dim >= 2:
            _eq("cleared", cx.cleared[:cx.T].cpu().numpy().astype(bool),
                np.isin(np.arange(cx.T) + cx.E, ref["h1_allpiv"]))


def test_tune_installs_and_caches(tmp_path):
    from flash_ph import autotune
    pts, t = cases.randn(3000, 10, 2), 1.9
    path = str(tmp_path / "tune.json")
    res = autotune.tune(pts, t, max_dim=2, repeats=2, persist=True, path=path)
    assert set(res) == {"screen_tile", "enum_block", "apparent_block", "queue_wgs"}
    for knob, r in res.items():
        assert r["value"] in autotune.candidates(knob)
        assert set(r["ms"]) == set(autotune.candidates(knob))
        assert all(v >= 0.0 for v in r["ms"].values())
        for size, d in r["buckets"]:
            assert autotune.query(knob, size, d) == r["value"]
    chosen = {k: r["value"] for k, r in res.items()}
    autotune.reset()
    assert autotune.load_cache(path) >= 3
    assert autotune.query("screen_tile", 3000, 10) == chosen["screen_tile"]
    # tuned settings still give the oracle's diagrams
    import oracle
    from flash_ph import rips_persistence
    ref = oracle.run(pts, t, 2)
    dg = rips_persistence(torch.from_numpy(pts).cuda(), t, max_dim=2)
    for k, name in enumerate(("h0", "h1", "h2")):
        _eq(name, dg[k].cpu().numpy(), ref[name])


def test_tune_skips_tile_for_matrix_core_screen():
    from flash_ph import autotune
    pts = cases.randn(4096, 32, 3)
    res = autotune.tune(pts, 5.0, max_dim=0, repeats=1)
    assert res == {}
