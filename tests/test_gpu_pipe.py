"""The reference-semantics reduction with many columns in flight
(fph_reduce_pipe) against the oracle and against the one-workgroup reducer
(fph_reduce_serial): diagrams bit for bit and the reference's number of
column additions (the oracle's xor1), with and without store overflow,
working bitmaps in LDS and in global memory, one cloud and batches, and the
heaviest cloud of the uneven-batch workload (n=4918, d=50: 38,164 residual
columns, 182,827 additions, 1,224 dropped columns)."""
from __future__ import annotations

import time

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


@pytest.fixture(autouse=True)
def _pipe_settings():
    from flash_ph import _pipeline
    old = (_pipeline.PIPE, _pipeline.PIPE_FORCE_GLOBAL, _pipeline.REPLAY,
           _pipeline._WAIT_SECONDS)
    _pipeline.REPLAY = False      # every cloud through the reference-order reducers
    _pipeline._WAIT_SECONDS = 30
    yield
    (_pipeline.PIPE, _pipeline.PIPE_FORCE_GLOBAL, _pipeline.REPLAY,
     _pipeline._WAIT_SECONDS) = old


def _uneven(i):
    return cases.batch_uneven(256, seed=0)[i]


CLOUDS = [("overflow_n1184_d10", lambda: _uneven(140), None),
          ("overflow_n1539_d50", lambda: _uneven(225), None),
          ("randn_n400_d3", lambda: cases.randn(400, 3, 2), 1.0),
          ("clifford_n600", lambda: cases.clifford(600, 3, 0.02), 0.6),
          ("randn_n2000_d10", lambda: cases.randn(2000, 10, 4), 2.2)]


@pytest.mark.parametrize("force_global", [False, True], ids=["lds", "global"])
@pytest.mark.parametrize("name,make,t", CLOUDS, ids=[c[0] for c in CLOUDS])
def test_pipe_matches_oracle(name, make, t, force_global):
    import oracle
    from flash_ph import _pipeline
    x = make()
    t = oracle.auto_threshold(x) if t is None else t
    ref = oracle.run(x, t, 1)
    _pipeline.PIPE, _pipeline.PIPE_FORCE_GLOBAL = True, force_global
    print(f"[gpu] pipe {name} global={force_global}", flush=True)
    dg, cx = _pipeline.compute(torch.from_numpy(x).cuda(), t, 1, semantics="reference")
    _eq("h0", dg[0].cpu().numpy(), ref["h0"])
    _eq("h1", dg[1].cpu().numpy(), ref["h1"])
    assert cx.info["h1_stats"]["xors"] == ref["stats"]["xor1"]


@pytest.mark.parametrize("mode", ["window32k", "nowindow"])
@pytest.mark.parametrize("name,make,t", CLOUDS, ids=[c[0] for c in CLOUDS])
def test_pipe_window(name, make, t, mode):
    """Bitmaps past the LDS arena (forced here): a window of rows in LDS over
    the global slot, capped at 32,768 rows so that it moves on these clouds
    (50K-2.8M rows), and the whole bitmap in the slot."""
    import oracle
    from flash_ph import _pipeline
    x = make()
    t = oracle.auto_threshold(x) if t is None else t
    ref = oracle.run(x, t, 1)
    old = _pipeline.PIPE_WINDOW, _pipeline.PIPE_WINDOW_UNITS
    _pipeline.PIPE, _pipeline.PIPE_FORCE_GLOBAL = True, True
    _pipeline.PIPE_WINDOW = mode != "nowindow"
    _pipeline.PIPE_WINDOW_UNITS = 1
    try:
        print(f"[gpu] pipe window {name} {mode}", flush=True)
        dg, cx = _pipeline.compute(torch.from_numpy(x).cuda(), t, 1, semantics="reference")
        _eq("h0", dg[0].cpu().numpy(), ref["h0"])
        _eq("h1", dg[1].cpu().numpy(), ref["h1"])
        assert cx.info["h1_stats"]["xors"] == ref["stats"]["xor1"]
    finally:
        _pipeline.PIPE_WINDOW, _pipeline.PIPE_WINDOW_UNITS = old


def test_pipe_window_batch_and_limits():
    """The window form on a batch (pipe = one-workgroup reducer = oracle) and
    with a lowered MAX_COL (the length bound's recount, the list-form
    handover)."""
    import oracle
    from flash_ph import _pipeline, rips_persistence_batch
    old = (_pipeline.PIPE_WINDOW, _pipeline.PIPE_WINDOW_UNITS, _pipeline.REF_MAX_COL)
    _pipeline.PIPE_FORCE_GLOBAL = True
    _pipeline.PIPE_WINDOW, _pipeline.PIPE_WINDOW_UNITS = True, 1
    try:
        allc = cases.batch_uneven(256, seed=0)
        clouds = [allc[i] for i in (140, 225, 3, 17, 40, 51)]
        ts = [float(oracle.auto_threshold(c)) for c in clouds]
        xs = [torch.from_numpy(c).cuda() for c in clouds]
        print("[gpu] pipe window batch", flush=True)
        _pipeline.PIPE = True
        got = [[d.cpu().numpy() for d in dg] for dg in rips_persistence_batch(xs, ts, 1)]
        for k in range(len(clouds)):
            ref = oracle.run(clouds[k], ts[k], 1)
            _eq(f"window batch cloud {k} h1", got[k][1], ref["h1"])
        x, t, max_col = cases.randn(700, 3, 3), 0.9, 24
        print("[gpu] pipe window max_col", flush=True)
        _pipeline.REF_MAX_COL = max_col
        ref = oracle.run(x, t, 1, max_col=max_col)
        assert ref["stats"]["truncations"] > 0
        dg, _ = _pipeline.compute(torch.from_numpy(x).cuda(), t, 1, semantics="reference")
        _eq("h1 truncated", dg[1].cpu().numpy(), ref["h1"])
    finally:
        _pipeline.PIPE_WINDOW, _pipeline.PIPE_WINDOW_UNITS, _pipeline.REF_MAX_COL = old


def test_pipe_heaviest_batch_cloud():
    """The uneven batch's heaviest cloud: the oracle's diagram and additions,
    and the one-workgroup reducer's diagram; both times printed."""
    import oracle
    from flash_ph import _pipeline
    x = _uneven(171)
    t = oracle.auto_threshold(x)
    ref = oracle.run(x, t, 1)
    assert ref["stats"]["pool_drops"] > 0 and ref["stats"]["xor1"] > 150_000
    xd = torch.from_numpy(x).cuda()
    out = {}
    for pipe in (True, False):
        _pipeline.PIPE = pipe
        print(f"[gpu] heaviest cloud pipe={pipe}", flush=True)
        _pipeline.compute(xd, t, 1, semantics="reference")
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        dg, cx = _pipeline.compute(xd, t, 1, semantics="reference")
        torch.cuda.synchronize()
        ms = 1000 * (time.perf_counter() - t0)
        print(f"[gpu] heaviest cloud pipe={pipe}: {ms:.1f} ms", flush=True)
        _eq(f"h1 pipe={pipe}", dg[1].cpu().numpy(), ref["h1"])
        assert cx.info["h1_stats"]["xors"] == ref["stats"]["xor1"]
        out[pipe] = dg[1].cpu().numpy()
    _eq("pipe vs serial", out[True], out[False])


@pytest.mark.parametrize("force_global", [False, True], ids=["lds", "global"])
def test_pipe_batch(force_global):
    """A batch of 24 uneven clouds (store overflow in some): pipe = the
    one-workgroup reducer = each cloud's oracle diagram."""
    import oracle
    from flash_ph import _pipeline, rips_persistence_batch
    idx = [140, 225, 3, 17, 40, 51, 66, 77, 88, 99, 110, 121, 132, 143, 154, 165, 176, 187,
           198, 209, 220, 231, 242, 253]
    allc = cases.batch_uneven(256, seed=0)
    clouds = [allc[i] for i in idx]
    ts = [float(oracle.auto_threshold(c)) for c in clouds]
    xs = [torch.from_numpy(c).cuda() for c in clouds]
    res = {}
    for pipe in (True, False):
        _pipeline.PIPE, _pipeline.PIPE_FORCE_GLOBAL = pipe, force_global and pipe
        print(f"[gpu] pipe batch pipe={pipe} global={force_global}", flush=True)
        res[pipe] = [[d.cpu().numpy() for d in dg] for dg in rips_persistence_batch(xs, ts, 1)]
    for a, b in zip(res[True], res[False]):
        _eq("batch h0", a[0], b[0])
        _eq("batch h1", a[1], b[1])
    for k in (0, 1, 5):
        ref = oracle.run(clouds[k], ts[k], 1)
        _eq(f"batch cloud {k} h1", res[True][k][1], ref["h1"])
