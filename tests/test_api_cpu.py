"""CPU-side checks of the drop-in boundary: the C-ABI library loads and
exports every entry point declared in include/flashph.h, argument
validation matches the reference's exceptions and messages, and the trivial
inputs (n = 0, 1) are answered without a GPU."""
from __future__ import annotations

import ctypes
import os
import re

import numpy as np
import pytest
import torch

import flash_ph
from flash_ph import RipsComplex, auto_threshold, enclosing_radius, rips_persistence
from flash_ph import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared_symbols():
    src = open(os.path.join(ROOT, "include", "flashph.h")).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(fph_[a-z0-9_]+)\s*\(", src)))


def test_library_exports_declared_symbols():
    lib = ctypes.CDLL(_lib.LIB_PATH)
    names = _declared_symbols()
    assert len(names) >= 20
    for name in names:
        assert hasattr(lib, name), name
        assert name in _lib.SIGNATURES, f"{name} not bound in flash_ph._lib"


def test_version_string():
    assert flash_ph.__version__ == "0.1.0"
    assert b"gfx950" in _lib.load().fph_version()


def test_workspace_queries_without_gpu():
    lib = _lib.load()
    assert lib.fph_boruvka_rounds(10_000) == 16
    assert lib.fph_boruvka_scratch_bytes(10_000) > 4 * 10_000
    assert lib.fph_reduce_scratch_bytes(1000) >= 20 * 1000


@pytest.mark.parametrize("bad", [-1.0, -1e-9])
def test_negative_threshold(bad):
    with pytest.raises(ValueError, match="max_edge_length must be non-negative"):
        rips_persistence(torch.zeros(4, 2), bad)
    with pytest.raises(ValueError, match="max_edge_length must be non-negative"):
        RipsComplex(torch.zeros(4, 2), bad)


def test_bad_shapes_and_dims():
    with pytest.raises(ValueError, match=r"points must have shape \(n, d\)"):
        rips_persistence(torch.zeros(4), 1.0)
    with pytest.raises(ValueError, match="max_dim must be 0, 1, or 2"):
        rips_persistence(torch.zeros(4, 2), 1.0, max_dim=3)
    with pytest.raises(ValueError, match="H2 requires n < 65536"):
        rips_persistence(torch.zeros(65536, 2), 1.0, max_dim=2)


def test_bad_backend():
    with pytest.raises(ValueError, match="backend must be one of"):
        rips_persistence(torch.zeros(4, 2), 1.0, max_dim=1, backend="cuda")


@pytest.mark.parametrize("max_dim", [0, 1, 2])
def test_empty_and_single(max_dim):
    d0 = rips_persistence(torch.empty(0, 3), 1.0, max_dim=max_dim)
    assert len(d0) == max_dim + 1 and all(d.shape == (0, 2) for d in d0)
    d1 = rips_persistence(torch.tensor([[1.0, 2.0, 3.0]]), 1.0, max_dim=max_dim)
    assert len(d1) == max_dim + 1
    assert d1[0].shape == (1, 2) and d1[0][0, 0] == 0 and d1[0][0, 1] == float("inf")
    assert all(d.shape == (0, 2) and d.dtype == torch.float32 for d in d1[1:])
    assert all(d.device.type == "cpu" for d in d0 + d1)


def test_numpy_single_point():
    d = rips_persistence(np.zeros((1, 4), np.float32), 1.0)
    assert d[0].shape == (1, 2)


@pytest.mark.skipif(torch.cuda.is_available(), reason="checks the no-GPU error")
def test_gpu_library_path_fails_loudly_without_a_gpu():
    """CUDA work needs the GPU: require_gpu raises where none is visible
    (there is no CPU stand-in for the HIP library)."""
    if torch.cuda.is_available():
        pytest.skip("a GPU is visible")
    with pytest.raises(RuntimeError, match="ROCm GPU"):
        _lib.require_gpu()
    from flash_ph import rips_persistence_batch
    with pytest.raises(RuntimeError, match="ROCm GPU"):
        rips_persistence_batch([torch.randn(10, 3)], 1.0)


def test_cpu_tensors_run_on_the_host_pipeline():
    """CPU tensors are computed on the CPU (as the reference's CPU route),
    by libflashph_cpu: the oracle's diagrams bit for bit, output on the
    CPU, no GPU needed."""
    import oracle
    pts = np.random.default_rng(0).standard_normal((80, 3)).astype(np.float32)
    dg = rips_persistence(torch.from_numpy(pts), 1.0, max_dim=2)
    ref = oracle.run(pts, 1.0, 2)["diagrams"]
    for a, b in zip(dg, ref):
        assert a.device.type == "cpu"
        assert np.array_equal(a.numpy().view(np.uint32), b.view(np.uint32))


def test_complex_errors():
    rc = RipsComplex(torch.zeros(3, 2), 1.0)
    with pytest.raises(RuntimeError, match="Call compute_persistence"):
        rc.betti_numbers()
    with pytest.raises(ValueError, match="Only Z/2Z coefficients"):
        rc.compute_persistence(homology_coeff_field=3)
    assert rc.num_vertices() == 3


def test_complex_single_point():
    rc = RipsComplex(torch.zeros(1, 2), 1.0)
    rc.compute_persistence(max_dim=1)
    assert rc.betti_numbers() == [1, 0]
    assert rc.persistence_intervals_in_dimension(1).shape == (0, 2)
    assert rc.persistence_intervals_in_dimension(5).shape == (0, 2)
    assert rc.persistent_betti_numbers(0.0, 1.0) == [1, 0]
    assert rc.persistence() == [(0, (0.0, float("inf")))]


def test_thresholds_on_cpu_tensors():
    """n <= 1 returns 0.0 (threshold.py:59-61, 106-108); CPU tensors run on
    the host pipeline with the filtration's FP32 arithmetic and torch's FP32
    quantile: the oracle's values bit for bit, k beyond 128 included."""
    import oracle
    pts = torch.from_numpy(np.random.default_rng(0).standard_normal((300, 3)).astype(np.float32))
    assert auto_threshold(pts[:1]) == 0.0
    assert auto_threshold(pts[:0]) == 0.0
    assert enclosing_radius(pts[:1]) == 0.0
    x = pts.numpy()
    for k, p in [(1, 50.0), (20, 95.0), (150, 90.0), (299, 100.0), (5000, 0.0)]:
        assert np.float32(auto_threshold(pts, k, p)) == np.float32(oracle.auto_threshold(x, k, p))
    assert np.float32(enclosing_radius(pts)) == np.float32(oracle.enclosing_radius(x))


def test_graph_argument_errors_without_gpu():
    from flash_ph.graph import RipsH0Graph
    with pytest.raises(ValueError):
        RipsH0Graph(100, 3, -1.0)
    with pytest.raises(ValueError):
        RipsH0Graph(1, 3, 1.0)
    with pytest.raises(ValueError):
        RipsH0Graph(92_683, 3, 1.0)
    with pytest.raises(ValueError):
        RipsH0Graph(100, 0, 1.0)
    with pytest.raises(RuntimeError):
        RipsH0Graph(100, 3, 1.0)


def test_complex_accessors_compare_in_fp32(monkeypatch):
    """RipsComplex's counts equal the reference's torch formulation
    (complex.py:96-130: FP32 tensors compared with Python scalars, i.e. in
    FP32), including bounds that are not FP32 numbers."""
    import flash_ph.complex as cm
    dg = [torch.tensor([[0.0, 0.1], [0.0, 0.30000001192092896], [0.0, float("inf")]]),
          torch.tensor([[0.1, 0.2], [0.2, float("inf")], [0.30000001192092896, 0.5]]),
          torch.empty(0, 2)]
    monkeypatch.setattr(cm, "rips_persistence", lambda *a, **k: [d.clone() for d in dg])
    rc = cm.RipsComplex(torch.zeros(5, 2), 1.0)
    rc.compute_persistence(max_dim=2)
    for lo, hi in [(0.1, 0.1), (0.3, 0.2), (0.0, 0.30000000000000004), (1.0, 0.0)]:
        want = [int(((d[:, 0] <= lo) & (d[:, 1] > hi)).sum()) if d.numel() else 0 for d in dg]
        assert rc.persistent_betti_numbers(lo, hi) == want, (lo, hi)
    assert rc.betti_numbers() == [int((d[:, 1] == float("inf")).sum()) for d in dg]
    got = rc.persistence(max_dim=2)
    want = sorted([(k, (float(b), float(e))) for k, d in enumerate(dg) for b, e in d.tolist()],
                  key=lambda r: (r[0], r[1][0], r[1][1]))
    assert got == want
    iv = rc.persistence_intervals_in_dimension(1)
    assert iv.dtype == np.float64 and np.array_equal(iv, dg[1].double().numpy())
    assert rc.persistence_intervals_in_dimension(2).shape == (0, 2)
    assert rc.persistence_intervals_in_dimension(-1).shape == (0, 2)
