"""GPU tests of the public API (mirrors the reference's
testing/test_persistence.py edge cases) and of pipeline invariants at the
full benchmark size."""
from __future__ import annotations

import numpy as np
import pytest
import torch

import cases

pytestmark = pytest.mark.gpu


def _u32(t):
    return t.detach().cpu().numpy().view(np.uint32)


def test_device_and_formats():
    from flash_ph import rips_persistence
    print("[gpu] api formats", flush=True)
    pts = torch.from_numpy(cases.randn(40, 3, 1)).cuda()
    for md in (0, 1, 2):
        dg = rips_persistence(pts, 1.2, max_dim=md)
        assert len(dg) == md + 1
        for d in dg:
            assert d.device.type == "cuda" and d.dtype == torch.float32 and d.shape[1] == 2


def test_numpy_and_cpu_inputs():
    from flash_ph import rips_persistence
    print("[gpu] api numpy/cpu inputs", flush=True)
    pts = cases.randn(50, 2, 3)
    a = rips_persistence(pts, 1.0)
    b = rips_persistence(torch.from_numpy(pts), 1.0)
    c = rips_persistence(torch.from_numpy(pts).cuda(), 1.0)
    assert a[0].device.type == "cuda" and b[0].device.type == "cpu"
    for x, y, z in zip(a, b, c):
        assert np.array_equal(_u32(x), _u32(y)) and np.array_equal(_u32(x), _u32(z))


def test_rips_filtration_and_boruvka_api():
    from flash_ph.rips import rips_filtration, _pack_sort_key
    from flash_ph.boruvka import gpu_boruvka_mst
    import oracle
    print("[gpu] api rips_filtration / gpu_boruvka_mst", flush=True)
    pts = cases.randn(300, 4, 8)
    f = rips_filtration(torch.from_numpy(pts).cuda(), 1.3)
    ref = oracle.run(pts, 1.3, 0)
    assert np.array_equal(f.edge_i.cpu().numpy(), ref["edge_i"])
    assert np.array_equal(f.vert_adj_rank.cpu().numpy(), ref["adj_rank"])
    key = _pack_sort_key(f.edge_dist_sq, f.edge_i, f.edge_j)
    assert bool((key[1:] > key[:-1]).all())
    E = f.edge_i.shape[0]
    mst, comp = gpu_boruvka_mst(f.edge_i, f.edge_j, torch.arange(E, device="cuda",
                                                                 dtype=torch.int32), 300)
    assert np.array_equal(mst.cpu().numpy(), ref["mst_idx"])
    assert np.array_equal(comp.cpu().numpy(), ref["comp"])


@pytest.mark.parametrize("n", [12_288, 12_289])
def test_boruvka_forms(n):
    """Forests of up to 12,288 vertices come from the one-workgroup LDS form,
    larger ones from the multi-launch form: both equal the oracle's (Kruskal
    by rank) and its component labels."""
    from flash_ph.rips import rips_filtration
    from flash_ph.boruvka import gpu_boruvka_mst
    import oracle
    print(f"[gpu] boruvka form n={n}", flush=True)
    pts = cases.randn(n, 3, 11)
    f = rips_filtration(torch.from_numpy(pts).cuda(), 0.12)
    ref = oracle.run(pts, 0.12, 0)
    E = f.edge_i.shape[0]
    mst, comp = gpu_boruvka_mst(f.edge_i, f.edge_j, torch.arange(E, device="cuda",
                                                                 dtype=torch.int32), n)
    assert 0 < len(ref["mst_idx"]) < n - 1  # several components
    assert np.array_equal(mst.cpu().numpy(), ref["mst_idx"])
    assert np.array_equal(comp.cpu().numpy(), ref["comp"])


def test_complex_api():
    from flash_ph import RipsComplex
    print("[gpu] api RipsComplex", flush=True)
    rc = RipsComplex(cases.circle(20), 0.5)
    rc.compute_persistence(max_dim=1)
    assert rc.betti_numbers() == [1, 1]
    iv = rc.persistence_intervals_in_dimension(0)
    assert iv.dtype == np.float64 and iv.shape == (20, 2)
    p = rc.persistence()
    assert p == sorted(p, key=lambda x: (x[0], x[1][0], x[1][1]))


def test_h1_standalone_entry():
    from flash_ph.rips import rips_filtration
    from flash_ph.boruvka import gpu_boruvka_mst
    from flash_ph.reduce_h1_gpu import rips_h1_gpu_native
    import oracle
    print("[gpu] api rips_h1_gpu_native", flush=True)
    pts = cases.randn(200, 3, 4)
    f = rips_filtration(torch.from_numpy(pts).cuda(), 0.9)
    E = f.edge_i.shape[0]
    mst, _ = gpu_boruvka_mst(f.edge_i, f.edge_j, torch.arange(E, device="cuda"), 200)
    h1 = rips_h1_gpu_native(f, 200, mst, 0.9, torch.device("cuda"), backend="gpu")
    ref = oracle.run(pts, 0.9, 1)["h1"]
    assert np.array_equal(_u32(h1), ref.view(np.uint32))


def test_deterministic_bits():
    from flash_ph import rips_persistence
    print("[gpu] determinism", flush=True)
    pts = torch.from_numpy(cases.randn(3000, 10, 5)).cuda()
    a = rips_persistence(pts, 2.0, max_dim=2)
    b = rips_persistence(pts, 2.0, max_dim=2)
    for x, y in zip(a, b):
        assert np.array_equal(_u32(x), _u32(y))


def test_n10000_d10_invariants():
    """n=10,000, d=10, t=1.6, H0-H2: every vertex accounted for in H0, every
    column either cleared, paired or essential, diagrams sorted by birth."""
    from flash_ph import _pipeline
    print("[gpu] n=10000 d=10 invariants", flush=True)
    pts = torch.from_numpy(cases.randn(10_000, 10, 0)).cuda()
    dg, cx = _pipeline.compute(pts, 1.6, 2)
    h0, h1, h2 = (d.cpu().numpy() for d in dg)
    assert h0.shape[0] == 10_000
    n_mst = int(cx.mst_mask[: cx.E].sum())
    assert int(np.isinf(h0[:, 1]).sum()) == 10_000 - n_mst
    for h in (h1, h2):
        assert np.all(np.diff(h[:, 0]) >= 0)
        fin = h[np.isfinite(h[:, 1])]
        assert np.all(fin[:, 1] > fin[:, 0])
    # Euler characteristic bookkeeping of the columns: E = forest + H1 columns
    kind1 = cx.kind1[: cx.E].cpu().numpy()
    assert int((kind1 == 0).sum()) == n_mst


def test_coo_export_and_giotto_route():
    """fph_coo_export: the sparse distance matrix the reference hands to
    giotto-ph (persistence.py:36-45) -- one (i, j, sqrt(dsq)) entry per edge
    plus the zero diagonal -- equal to the oracle's edges; its minimum
    spanning forest gives the H0 deaths.  Without giotto-ph installed,
    backend='giotto' raises ModuleNotFoundError as the reference does."""
    import importlib.util

    import oracle
    from scipy.sparse.csgraph import minimum_spanning_tree

    from flash_ph import rips_persistence
    from flash_ph.persistence import coo_distance_matrix
    from flash_ph.rips import rips_filtration
    print("[gpu] api coo export", flush=True)
    pts = cases.randn(400, 3, 8)
    ref = oracle.run(pts, 0.8, 0)
    f = rips_filtration(torch.from_numpy(pts).cuda(), 0.8)
    dm = coo_distance_matrix(f)
    E = ref["stats"]["E"]
    assert dm.shape == (400, 400) and dm.nnz == E + 400
    assert np.array_equal(dm.row[:E], ref["edge_i"]) and np.array_equal(dm.col[:E], ref["edge_j"])
    assert np.array_equal(dm.data[:E].view(np.uint32),
                          oracle.sqrt_f32(ref["edge_dsq"]).view(np.uint32))
    assert np.array_equal(dm.row[E:], np.arange(400)) and not dm.data[E:].any()
    fin = np.sort(minimum_spanning_tree(dm.tocsr()).data)
    h0 = ref["h0"][np.isfinite(ref["h0"][:, 1]), 1]
    assert np.array_equal(fin.astype(np.float32), h0)
    if importlib.util.find_spec("gph") is None:
        with pytest.raises(ModuleNotFoundError):
            rips_persistence(torch.from_numpy(pts).cuda(), 0.8, max_dim=1, backend="giotto")
