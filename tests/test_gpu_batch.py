"""GPU parity of the batched path: per-cloud kNN thresholds and diagrams of
a batch of uneven clouds, bit-exact against the oracle run cloud by cloud
and against the plain single-cloud path."""
from __future__ import annotations

import numpy as np
import pytest
import torch

import cases
import oracle

pytestmark = pytest.mark.gpu


def _u32(a):
    if isinstance(a, torch.Tensor):
        a = a.detach().cpu().numpy()
    return np.ascontiguousarray(a, dtype=np.float32).view(np.uint32)


def _mixed_clouds():
    """Uneven sizes and dimensions, including the degenerate ones."""
    out = [
        cases.randn(0, 3, 1),
        cases.randn(1, 3, 2),
        cases.randn(2, 3, 3),
        cases.randn(37, 1, 4),
        cases.randn(65, 3, 5),
        cases.sphere(130, 4, 6, 0.05),
        cases.randn(64, 10, 7),
        cases.sphere(250, 10, 8, 0.05),
        cases.randn(90, 17, 9),
        cases.sphere(300, 50, 10, 0.05),
        cases.randn(70, 70, 11),
        cases.clifford(200, 12, 0.02),
        cases.grid3(),
        cases.circle(30),
    ]
    return out


@pytest.mark.parametrize("k", [1, 5, 20, 33, 100, 128, 129, 200, 1000])
def test_knn_threshold_batch_vs_oracle(k):
    from flash_ph import _pipeline
    print(f"[gpu] knn batch k={k}", flush=True)
    clouds = [c for c in _mixed_clouds()]
    pts = [torch.from_numpy(c).cuda() for c in clouds]
    L = _pipeline.layout_batch(pts, pts[0].device)
    t, thr_sq, kth = _pipeline.knn_threshold_batch(L, k, 95.0, want_kth=True)
    t = t.cpu().numpy()
    thr_sq = thr_sq.cpu().numpy()
    kth = kth.cpu().numpy()
    vb = L.vbase_host
    for c, x in enumerate(clouds):
        n = x.shape[0]
        if n <= 1:
            assert t[c] == 0.0
            continue
        kk = max(1, min(k, n - 1))
        ref_kth = oracle.sqrt_f32(oracle.knn_kth_dsq(x, kk))
        assert np.array_equal(_u32(kth[vb[c]:vb[c + 1]]), _u32(ref_kth)), f"cloud {c}"
        ref_t = np.float32(oracle.auto_threshold(x, k, 95.0))
        assert _u32(t[c:c + 1])[0] == _u32(np.array([ref_t]))[0], f"cloud {c}"
        assert thr_sq[c] == oracle.threshold_sq_f32(float(t[c]))


def test_auto_threshold_batch_percentiles():
    from flash_ph import auto_threshold_batch
    print("[gpu] auto_threshold_batch percentiles", flush=True)
    clouds = [cases.randn(200, 5, 1), cases.sphere(300, 3, 2, 0.05)]
    for p in (0.0, 12.5, 50.0, 95.0, 100.0, 150.0, -3.0):
        got = auto_threshold_batch(clouds, k=7, percentile=p)
        for c, x in enumerate(clouds):
            assert np.float32(got[c]) == np.float32(oracle.auto_threshold(x, 7, p))


def _check_against_oracle(clouds, ts, dgms, max_dim):
    for c, (x, t) in enumerate(zip(clouds, ts)):
        ref = oracle.run(x, t, max_dim)["diagrams"]
        assert len(dgms[c]) == max_dim + 1
        for dim in range(max_dim + 1):
            got = dgms[c][dim]
            assert got.shape == ref[dim].shape, (c, dim, got.shape, ref[dim].shape)
            assert np.array_equal(_u32(got), _u32(ref[dim])), (c, dim)


@pytest.mark.parametrize("max_dim", [0, 1])
def test_batch_fixed_thresholds_vs_oracle(max_dim):
    from flash_ph import rips_persistence_batch
    print(f"[gpu] batch fixed thresholds max_dim={max_dim}", flush=True)
    clouds = _mixed_clouds()
    ts = [1.0, 1.0, 5.0, 0.5, 1.2, 0.9, 3.5, 0.8, 5.0, 1.0, 11.0, 0.6, 2.0, 1.5]
    dgms = rips_persistence_batch(clouds, ts, max_dim=max_dim)
    _check_against_oracle(clouds, ts, dgms, max_dim)


def test_batch_auto_thresholds_vs_oracle():
    from flash_ph import rips_persistence_batch, auto_threshold_batch
    print("[gpu] batch auto thresholds", flush=True)
    clouds = cases.batch_uneven(10, seed=3, nmin=20, nmax=600)
    ts = auto_threshold_batch(clouds)
    dgms = rips_persistence_batch(clouds)
    _check_against_oracle(clouds, ts, dgms, 1)


def test_batch_single_threshold_and_devices():
    from flash_ph import rips_persistence_batch
    print("[gpu] batch scalar threshold / cpu tensors", flush=True)
    clouds = [cases.randn(80, 3, s) for s in range(5)]
    a = rips_persistence_batch(clouds, 1.1)
    b = rips_persistence_batch([torch.from_numpy(c) for c in clouds], [1.1] * 5)
    assert a[0][0].device.type == "cuda" and b[0][0].device.type == "cpu"
    for x, y in zip(a, b):
        for u, v in zip(x, y):
            assert np.array_equal(_u32(u), _u32(v))
    _check_against_oracle(clouds, [1.1] * 5, a, 1)


def test_batch_errors():
    from flash_ph import rips_persistence_batch, auto_threshold_batch
    with pytest.raises(ValueError):
        rips_persistence_batch([cases.randn(10, 2, 0)], 1.0, max_dim=2)
    with pytest.raises(ValueError):
        rips_persistence_batch([cases.randn(10, 2, 0)], [1.0, 2.0])
    with pytest.raises(ValueError):
        rips_persistence_batch([cases.randn(10, 2, 0)], -1.0)
    assert rips_persistence_batch([], 1.0) == []


def test_batch_uneven_full_size_vs_plain_path():
    """The benchmark's batch (256 clouds, n log-uniform in [200, 5000],
    d in {4, 10, 50}, auto thresholds): every cloud's diagrams equal the
    plain path's for the same threshold."""
    from flash_ph import rips_persistence, rips_persistence_batch, auto_threshold_batch
    print("[gpu] batch 256 uneven clouds vs plain path", flush=True)
    clouds = cases.batch_uneven(256, seed=0)
    ts = auto_threshold_batch(clouds)
    dgms = rips_persistence_batch(clouds, ts)
    for c, (x, t) in enumerate(zip(clouds, ts)):
        ref = rips_persistence(torch.from_numpy(x).cuda(), t, max_dim=1)
        for dim in range(2):
            assert np.array_equal(_u32(dgms[c][dim]), _u32(ref[dim])), (c, dim)
        # H0: one row per point; essential rows = components
        assert dgms[c][0].shape[0] == x.shape[0]


def test_auto_threshold_and_enclosing_radius_api():
    """The drop-in single-cloud functions run the same kernels: bit-exact
    against the oracle for k <= 128; the k > 128 path keeps the reference's
    cdist formulation and agrees to cdist rounding."""
    from flash_ph import auto_threshold, enclosing_radius
    print("[gpu] auto_threshold / enclosing_radius api", flush=True)
    for seed, (n, d) in enumerate([(300, 3), (500, 10), (257, 50), (129, 80)]):
        x = cases.randn(n, d, 50 + seed)
        for k, p in ((20, 95.0), (1, 50.0), (128, 99.0)):
            got = auto_threshold(torch.from_numpy(x).cuda(), k=k, percentile=p)
            assert np.float32(got) == np.float32(oracle.auto_threshold(x, k, p)), (n, d, k)
        assert np.float32(auto_threshold(x)) == np.float32(oracle.auto_threshold(x))
        assert np.float32(enclosing_radius(x)) == np.float32(oracle.enclosing_radius(x))
        assert np.float32(enclosing_radius(torch.from_numpy(x))) == np.float32(
            oracle.enclosing_radius(x))
    x = cases.randn(400, 4, 9)
    for k in (129, 200, 399, 5000):   # the radix-select path, bit-exact
        big_k = auto_threshold(x, k=k, percentile=90.0)
        assert np.float32(big_k) == np.float32(oracle.auto_threshold(x, k, 90.0)), k
    r_all = enclosing_radius(x)
    r_sub = enclosing_radius(x, subsample=100)
    assert 0 < r_sub and abs(r_sub - r_all) / r_all < 0.5
    assert auto_threshold(x, subsample=50) > 0


def test_enclosing_radius_batch_kernel():
    from flash_ph import _pipeline
    from flash_ph.threshold import _enclosing_radius_batch
    print("[gpu] enclosing radius batch", flush=True)
    clouds = _mixed_clouds()
    pts = [torch.from_numpy(c).cuda() for c in clouds]
    L = _pipeline.layout_batch(pts, pts[0].device)
    r = _enclosing_radius_batch(L).cpu().numpy()
    for c, x in enumerate(clouds):
        assert np.float32(r[c]) == np.float32(oracle.enclosing_radius(x)), c
