"""GPU parity: every stage of the HIP pipeline against the CPU oracle, bit
for bit (integer, index and FP32 arrays compared with array_equal on the raw
bits)."""
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


def _check_case(pts, t, max_dim, semantics="reference"):
    import oracle
    from flash_ph import _pipeline

    ref = oracle.run(pts, t, max_dim, exact=(semantics == "exact"))
    x = torch.from_numpy(pts).cuda()
    dgms, cx = _pipeline.compute(x, t, max_dim, semantics=semantics)
    torch.cuda.synchronize()
    E = cx.E
    # stage (a): edges and CSR
    _eq("edge_i", _cpu(cx.edge_i), ref["edge_i"])
    _eq("edge_j", _cpu(cx.edge_j), ref["edge_j"])
    _eq("edge_dsq", _cpu(cx.edge_dsq), ref["edge_dsq"])
    _eq("adj_ptr", _cpu(cx.adj_ptr), ref["adj_ptr"])
    _eq("adj_idx", _cpu(cx.adj_idx), ref["adj_idx"])
    _eq("adj_rank", _cpu(cx.adj_rank), ref["adj_rank"])
    _eq("adj_dsq", _cpu(cx.adj_dsq), ref["adj_dsq"])
    # stage (b): forest and H0
    mst = np.nonzero(_cpu(cx.mst_mask)[:E])[0]
    _eq("mst_idx", mst, ref["mst_idx"])
    _eq("mst_kruskal", mst, np.nonzero(ref["mst_kruskal"])[0])
    _eq("h0", _cpu(dgms[0]), ref["h0"])
    if max_dim < 1:
        return
    # stage (d): triangles in filtration order
    T = cx.T
    assert T == ref["stats"]["T"]
    if T:
        te = _cpu(cx.tri_e)[:T]
        ours = np.sort(np.stack([_cpu(cx.edge_i)[te], _cpu(cx.edge_j)[te], _cpu(cx.tri_w)[:T]], 1),
                       axis=1)
        order = np.argsort(ref["tri_global_rank"])
        theirs = np.stack([ref["tri_v0"], ref["tri_v1"], ref["tri_v2"]], 1)[order]
        _eq("triangles_by_rank", ours, theirs)
        # cofacet columns of non-forest edges
        off, dat = _cpu(cx.cof_off), _cpu(cx.cof_data).view(np.uint32)
        roff, rdat = ref["cof_off"], ref["cof_data"]
        mask = _cpu(cx.mst_mask)[:E]
        for e in range(E):
            if mask[e]:
                assert off[e + 1] == off[e]
                continue
            _eq(f"cof[{e}]", dat[off[e]:off[e + 1]].astype(np.int64),
                rdat[roff[e]:roff[e + 1]].astype(np.int64))
        # stage (c): apparent pairs
        _eq("apparent1", (_cpu(cx.kind1)[:E] == 2).astype(np.uint8), ref["apparent1"])
    # H1 diagram, row order included
    _eq("h1", _cpu(dgms[1]), ref["h1"])
    if max_dim < 2:
        return
    Q = cx.Q
    assert Q == ref["stats"]["Q"]
    if Q:
        qe = _cpu(cx.tet_e)[:Q]
        key = _cpu(cx.tet_key)[:Q].view(np.uint64)
        ours = np.sort(np.stack([_cpu(cx.edge_i)[qe], _cpu(cx.edge_j)[qe],
                                 (key >> np.uint64(32)).astype(np.int32),
                                 (key & np.uint64(0xFFFFFFFF)).astype(np.int32)], 1), axis=1)
        order = np.argsort(ref["tet_global_rank"])
        theirs = np.stack([ref["tet_v0"], ref["tet_v1"], ref["tet_v2"], ref["tet_v3"]], 1)[order]
        _eq("tetrahedra_by_rank", ours, theirs)
        _eq("apparent2", (_cpu(cx.kind2)[:T] == 2).astype(np.uint8), ref["apparent2"])
    _eq("h2", _cpu(dgms[2]), ref["h2"])


@pytest.mark.parametrize("name,make,t,max_dim", cases.SMALL, ids=[c[0] for c in cases.SMALL])
def test_stage_parity_small(name, make, t, max_dim):
    print(f"[gpu] parity {name}", flush=True)
    _check_case(make(), t, max_dim)


@pytest.mark.parametrize("name,make,t,max_dim", cases.MEDIUM, ids=[c[0] for c in cases.MEDIUM])
def test_stage_parity_medium(name, make, t, max_dim):
    print(f"[gpu] parity {name}", flush=True)
    _check_case(make(), t, max_dim)


H1_CASES = [c for c in cases.SMALL + cases.MEDIUM if c[3] == 1]


@pytest.mark.parametrize("name,make,t,max_dim", H1_CASES, ids=[c[0] for c in H1_CASES])
def test_exact_semantics_h1(name, make, t, max_dim):
    """The parallel complete reduction (backend='exact') against the oracle's
    complete reduction."""
    print(f"[gpu] exact semantics {name}", flush=True)
    _check_case(make(), t, 1, semantics="exact")


def _store_overflow_clouds():
    """Clouds of the uneven-batch workload whose reference store overflows
    (the oracle reports dropped columns): n=1184 d=10 and n=1539 d=50."""
    clouds = cases.batch_uneven(256, seed=0)
    return [clouds[140], clouds[225]]


@pytest.mark.parametrize("idx", [0, 1])
def test_store_overflow_reference_and_exact(idx):
    """Where the reference's capacity-limited store drops columns, backend
    'auto' reproduces the reference bit for bit and backend 'exact' the
    complete reduction."""
    import oracle
    from flash_ph import rips_persistence
    x = _store_overflow_clouds()[idx]
    t = oracle.auto_threshold(x)
    print(f"[gpu] store overflow case {idx} n={x.shape[0]} d={x.shape[1]}", flush=True)
    ref = oracle.run(x, t, 1)
    assert ref["stats"]["pool_drops"] > 0
    got = rips_persistence(torch.from_numpy(x).cuda(), t, max_dim=1)
    _eq("h0", _cpu(got[0]), ref["h0"])
    _eq("h1 reference", _cpu(got[1]), ref["h1"])
    ex = oracle.run(x, t, 1, exact=True)
    assert ex["stats"]["pool_drops"] == 0
    got_ex = rips_persistence(torch.from_numpy(x).cuda(), t, max_dim=1, backend="exact")
    _eq("h1 exact", _cpu(got_ex[1]), ex["h1"])


_REPLAY_CASES = {
    "randn_n300_d4": lambda: (cases.randn(300, 4, 8), 1.3),
    "clifford_n600": lambda: (cases.clifford(600, 3, 0.02), 0.6),
    "randn_n10000_d10": lambda: (cases.randn(10_000, 10, 0), 1.6),
}


@pytest.mark.parametrize("name", list(_REPLAY_CASES))
def test_replay_counts_the_reference_store(name):
    """Reference semantics, one cloud whose reference store drops nothing:
    the parallel reduction is kept, and its replay of the reference's
    additions counts exactly the reference's additions, stored entries and
    stored columns (oracle xor1, store1, stored1)."""
    import oracle
    from flash_ph import _pipeline
    pts, t = _REPLAY_CASES[name]()
    print(f"[gpu] replay {name}", flush=True)
    ref = oracle.run(pts, t, 1)
    assert ref["stats"]["pool_drops"] == 0 and ref["stats"]["truncations"] == 0
    dg, cx = _pipeline.compute(torch.from_numpy(pts).cuda(), t, 1, semantics="reference")
    assert cx.info["h1_stats"].get("replay_ok") is True
    rp = cx.info["h1_replay"]
    assert rp["status"] == 0
    assert rp["xors"] == ref["stats"]["xor1"]
    assert rp["held"] == ref["stats"]["store1"]
    assert rp["held_columns"] == ref["stats"]["stored1"]
    _eq("h0", _cpu(dg[0]), ref["h0"])
    _eq("h1", _cpu(dg[1]), ref["h1"])


def test_replay_detects_store_overflow():
    """Where the reference's store drops columns, the replay rejects the
    parallel pairs and the reference's own reducer runs (bit-exact); a cloud
    whose rows exceed a wavefront's bitmap goes to that reducer directly."""
    import oracle
    from flash_ph import _pipeline
    for idx in (0, 1):
        x = _store_overflow_clouds()[idx]
        t = oracle.auto_threshold(x)
        print(f"[gpu] replay overflow case {idx} n={x.shape[0]} d={x.shape[1]}", flush=True)
        ref = oracle.run(x, t, 1)
        assert ref["stats"]["pool_drops"] > 0
        dg, cx = _pipeline.compute(torch.from_numpy(x).cuda(), t, 1, semantics="reference")
        assert "replay_ok" not in cx.info["h1_stats"]  # the serial reducer's stats
        if ref["stats"]["T"] <= _pipeline._REPLAY_MAX_ROWS:
            rp = cx.info["h1_replay"]
            assert rp["status"] == 1 or rp["held"] > rp["cap"]
        else:
            assert "h1_replay" not in cx.info
        _eq("h1", _cpu(dg[1]), ref["h1"])


def test_serial_global_bitmap():
    """The serial reducer with its working-column bitmap in global memory
    (clouds with more rows than LDS holds, as the 2,000-point Clifford
    batch; forced here by a small LDS budget) gives the oracle's
    reference-semantics diagrams, one cloud and a batch, with and without
    store overflow."""
    import oracle
    from flash_ph import _pipeline, rips_persistence_batch
    old = _pipeline._BITMAP_LDS_WORDS, _pipeline.REPLAY
    _pipeline._BITMAP_LDS_WORDS = 64
    _pipeline.REPLAY = False
    try:
        clouds = [(_store_overflow_clouds()[0], None), (cases.randn(400, 3, 2), 1.0),
                  (cases.clifford(600, 3, 0.02), 0.6)]
        refs = []
        for x, t in clouds:
            t = oracle.auto_threshold(x) if t is None else t
            print(f"[gpu] serial global bitmap n={x.shape[0]} d={x.shape[1]}", flush=True)
            ref = oracle.run(x, t, 1)
            refs.append((t, ref))
            dg, cx = _pipeline.compute(torch.from_numpy(x).cuda(), t, 1, semantics="reference")
            _eq("h1", _cpu(dg[1]), ref["h1"])
        print("[gpu] serial global bitmap batch", flush=True)
        got = rips_persistence_batch([torch.from_numpy(x).cuda() for x, _ in clouds],
                                     [t for t, _ in refs], max_dim=1)
        for (t, ref), g in zip(refs, got):
            _eq("batch h1", _cpu(g[1]), ref["h1"])
    finally:
        _pipeline._BITMAP_LDS_WORDS, _pipeline.REPLAY = old


@pytest.mark.parametrize("units", [1, 0, -1], ids=["window32k", "window", "nowindow"])
def test_serial_window(units):
    """The serial reducer's global bitmap with a window of rows in LDS: a
    32,768-row window that moves several times per column on clouds of
    50K-72K rows, the largest window, and no window.  Reference semantics
    (with and without store overflow, one cloud and a batch), a lowered
    MAX_COL (the length bound's recount and the list-form handover) and the
    complete reduction in serial order, each against the oracle."""
    import oracle
    from flash_ph import _pipeline, rips_persistence_batch
    old = (_pipeline._BITMAP_LDS_WORDS, _pipeline.REPLAY, _pipeline.SERIAL_WINDOW,
           _pipeline.SERIAL_WINDOW_UNITS, _pipeline.REF_MAX_COL, _pipeline.EXACT_SERIAL_MIN_ROWS)
    _pipeline._BITMAP_LDS_WORDS = 64
    _pipeline.REPLAY = False
    _pipeline.SERIAL_WINDOW = units >= 0
    _pipeline.SERIAL_WINDOW_UNITS = max(units, 0)
    try:
        clouds = [(_store_overflow_clouds()[0], None), (cases.randn(400, 3, 2), 1.0),
                  (cases.clifford(600, 3, 0.02), 0.6)]
        refs = []
        for x, t in clouds:
            t = oracle.auto_threshold(x) if t is None else t
            print(f"[gpu] serial window units={units} n={x.shape[0]} d={x.shape[1]}", flush=True)
            ref = oracle.run(x, t, 1)
            refs.append((t, ref))
            dg, cx = _pipeline.compute(torch.from_numpy(x).cuda(), t, 1, semantics="reference")
            _eq("h1", _cpu(dg[1]), ref["h1"])
            assert cx.info["h1_stats"]["xors"] == ref["stats"]["xor1"]
        print(f"[gpu] serial window units={units} batch", flush=True)
        got = rips_persistence_batch([torch.from_numpy(x).cuda() for x, _ in clouds],
                                     [t for t, _ in refs], max_dim=1)
        for (t, ref), g in zip(refs, got):
            _eq("batch h1", _cpu(g[1]), ref["h1"])
        x, t, max_col = cases.randn(700, 3, 3), 0.9, 24
        print(f"[gpu] serial window units={units} max_col={max_col}", flush=True)
        _pipeline.REF_MAX_COL = max_col
        ref = oracle.run(x, t, 1, max_col=max_col)
        assert ref["stats"]["truncations"] > 0
        dg, _ = _pipeline.compute(torch.from_numpy(x).cuda(), t, 1, semantics="reference")
        _eq("h1 truncated", _cpu(dg[1]), ref["h1"])
        _pipeline.REF_MAX_COL = old[4]
        print(f"[gpu] serial window units={units} complete reduction", flush=True)
        _pipeline.EXACT_SERIAL_MIN_ROWS = 0
        x = clouds[0][0]
        t = refs[0][0]
        ex = oracle.run(x, t, 1, exact=True)
        dg, _ = _pipeline.compute(torch.from_numpy(x).cuda(), t, 1, semantics="exact")
        _eq("h1 exact", _cpu(dg[1]), ex["h1"])
    finally:
        (_pipeline._BITMAP_LDS_WORDS, _pipeline.REPLAY, _pipeline.SERIAL_WINDOW,
         _pipeline.SERIAL_WINDOW_UNITS, _pipeline.REF_MAX_COL,
         _pipeline.EXACT_SERIAL_MIN_ROWS) = old


@pytest.mark.parametrize("variant", [0, 1])
def test_serial_reducer_variants(variant):
    """Both forms of the working column (bitmap, sorted list) of the
    reference-semantics reducer give the reference's diagrams, on clouds
    with and without store overflow."""
    import oracle
    from flash_ph import _pipeline
    print(f"[gpu] serial variant {variant}", flush=True)
    old = _pipeline.SERIAL_VARIANT, _pipeline.REPLAY
    _pipeline.SERIAL_VARIANT = variant
    _pipeline.REPLAY = False  # every cloud through the serial reducer
    try:
        for x, t in [(_store_overflow_clouds()[0], None), (cases.randn(400, 3, 2), 1.0),
                     (cases.clifford(600, 3, 0.02), 0.6)]:
            t = oracle.auto_threshold(x) if t is None else t
            ref = oracle.run(x, t, 1)
            dg, cx = _pipeline.compute(torch.from_numpy(x).cuda(), t, 1, semantics="reference")
            _eq("h1", _cpu(dg[1]), ref["h1"])
    finally:
        _pipeline.SERIAL_VARIANT, _pipeline.REPLAY = old


def test_exact_serial_and_parallel_agree():
    """The complete reduction run in serial order (dense complexes) and in
    parallel (sparse ones) give the oracle's complete reduction."""
    import oracle
    from flash_ph import _pipeline
    print("[gpu] exact serial vs parallel", flush=True)
    x = _store_overflow_clouds()[0]
    t = oracle.auto_threshold(x)
    ex = oracle.run(x, t, 1, exact=True)
    old = _pipeline.EXACT_SERIAL_MIN_ROWS
    try:
        for thr in (0, 1 << 40):
            _pipeline.EXACT_SERIAL_MIN_ROWS = thr
            dg, cx = _pipeline.compute(torch.from_numpy(x).cuda(), t, 1, semantics="exact")
            _eq(f"h1 exact (serial threshold {thr})", _cpu(dg[1]), ex["h1"])
        _pipeline.EXACT_SERIAL_MIN_ROWS = 0   # H1 and H2 of max_dim=2 in serial order
        for name, make, tt, md in cases.SMALL[:8]:
            if md == 2:
                _check_case(make(), tt, 2)
    finally:
        _pipeline.EXACT_SERIAL_MIN_ROWS = old


def test_device_queue_list_form():
    """n=10K, d=10 at t=1.72: 94,249 triangle rows do not fit a wavefront's
    bitmap arena (2496 words), so the device-wide queue runs phase A and
    sorted-list working columns; the H1 diagram is the oracle's complete
    reduction, bit for bit."""
    import oracle
    from flash_ph import _pipeline
    print("[gpu] device queue, list working columns", flush=True)
    x = cases.randn(10000, 10, 0)
    t = 1.72
    ex = oracle.run(x, t, 1, exact=True)
    dg, cx = _pipeline.compute(torch.from_numpy(x).cuda(), t, 1, semantics="exact")
    assert cx.T == ex["stats"]["T"] and (cx.T + 31) // 32 > 2496 and cx.T <= 100_000
    _eq("h1 (list-form device queue)", _cpu(dg[1]), ex["h1"])
