"""ctypes front end of the CPU oracle (fph_oracle.c).

Test infrastructure only: tests/, __graft_entry__.smoke() and bench.py's
cpu_baseline leg use it as the checker.  The flash_ph package never imports it.

``run(points, max_edge_length, max_dim)`` returns every intermediate array of
the restated reference pipeline (edges, CSR, MST, triangles, cofacets,
apparent pairs, pairs, diagrams) as numpy arrays keyed by name, plus the
diagrams in the reference's output format (list of (K, 2) float32 arrays).
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle.so")

# Array ids: keep in the order of the enum in fph_oracle.c.
NAMES = [
    "edge_i", "edge_j", "edge_dsq",
    "adj_ptr", "adj_idx", "adj_rank", "adj_dsq",
    "mst_idx", "comp", "mst_kruskal", "h0",
    "tri_v0", "tri_v1", "tri_v2", "tri_max_rank", "tri_global_rank",
    "cof_off", "cof_data", "cmf1", "apparent1",
    "h1_pb", "h1_pd", "h1_eb", "h1_allpiv", "h1",
    "tet_v0", "tet_v1", "tet_v2", "tet_v3", "tet_max_rank", "tet_global_rank",
    "cof2_off", "cof2_data", "apparent2",
    "h2_pb", "h2_pd", "h2_eb", "h2",
    "stats",
]
STATS = ["E", "T", "Q", "boruvka_rounds", "residual1", "residual2",
         "pool_drops", "truncations", "xor1", "xor2", "ncomp", "store1", "stored1"]

_DTYPES = {
    "edge_dsq": np.float32, "adj_dsq": np.float32,
    "mst_idx": np.int64, "comp": np.int64, "mst_kruskal": np.uint8,
    "apparent1": np.uint8, "apparent2": np.uint8,
    "h0": np.float32, "h1": np.float32, "h2": np.float32,
    "stats": np.int64,
}

_lib = None


def build(force: bool = False) -> str:
    """Compile liboracle.so with the committed Makefile."""
    if force or not os.path.exists(_LIB_PATH):
        subprocess.run(["make", "-B", "-s", "-C", _HERE], check=True)
    return _LIB_PATH


def _load():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        lib = ctypes.CDLL(_LIB_PATH)
        lib.orc_run.restype = ctypes.c_void_p
        lib.orc_run.argtypes = [ctypes.c_void_p, ctypes.c_int64, ctypes.c_int,
                                ctypes.c_float, ctypes.c_int, ctypes.c_int]
        lib.orc_run_ex.restype = ctypes.c_void_p
        lib.orc_run_ex.argtypes = [ctypes.c_void_p, ctypes.c_int64, ctypes.c_int,
                                   ctypes.c_float, ctypes.c_int, ctypes.c_int,
                                   ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p,
                                   ctypes.c_int64, ctypes.c_int64, ctypes.c_int64]
        lib.orc_vertex_neighbours.restype = ctypes.c_int64
        lib.orc_vertex_neighbours.argtypes = [ctypes.c_void_p, ctypes.c_int64, ctypes.c_int,
                                              ctypes.c_float, ctypes.c_int64, ctypes.c_void_p,
                                              ctypes.c_void_p, ctypes.c_int64]
        lib.orc_len.restype = ctypes.c_int64
        lib.orc_len.argtypes = [ctypes.c_void_p, ctypes.c_int]
        lib.orc_get.restype = ctypes.c_int
        lib.orc_get.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p]
        lib.orc_free.restype = None
        lib.orc_free.argtypes = [ctypes.c_void_p]
        lib.orc_dist_sq.restype = ctypes.c_float
        lib.orc_dist_sq.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_int64,
                                    ctypes.c_int64]
        lib.orc_fmaf.restype = ctypes.c_float
        lib.orc_far_dsq.restype = ctypes.c_int
        lib.orc_far_dsq.argtypes = [ctypes.c_void_p, ctypes.c_int64, ctypes.c_int,
                                    ctypes.c_void_p]
        lib.orc_knn_kth.restype = ctypes.c_int
        lib.orc_knn_kth.argtypes = [ctypes.c_void_p, ctypes.c_int64, ctypes.c_int, ctypes.c_int,
                                    ctypes.c_void_p]
        lib.orc_fmaf.argtypes = [ctypes.c_float, ctypes.c_float, ctypes.c_float]
        _lib = lib
    return _lib


def threshold_sq_f32(max_edge_length: float) -> np.float32:
    """t*t in Python double, rounded to FP32 (rips.py:125 as a Triton fp32 arg)."""
    return np.float32(float(max_edge_length) * float(max_edge_length))


def run(points, max_edge_length: float, max_dim: int = 1, exact: bool = False,
        edges=None, max_col: int = 0, max_pool: int = 0) -> dict:
    """exact=False: the reference's results (its capacity-limited H1 store
    for max_dim <= 1; the complete reduction giotto-ph performs for
    max_dim = 2).  exact=True: the complete reduction everywhere.

    edges: optional (edge_i, edge_j, edge_dsq) arrays in any order -- the
    pipeline then starts from this edge set (put into the reference's
    emission order, then key-sorted) instead of enumerating all pairs, for
    clouds too large for the O(n^2 d) enumeration.  max_col / max_pool: the
    reference's MAX_COL and store-entry cap (_numba_utils.py:10-11; 0 = its
    values), lowered by tests to reach the truncation and drop branches."""
    lib = _load()
    pts = np.ascontiguousarray(np.asarray(points, dtype=np.float32))
    if pts.ndim != 2:
        raise ValueError("points must be (n, d)")
    n, d = pts.shape
    out: dict = {}
    if n <= 1:
        h0 = (np.array([[0.0, np.inf]], dtype=np.float32) if n == 1
              else np.empty((0, 2), np.float32))
        out["diagrams"] = [h0] + [np.empty((0, 2), np.float32) for _ in range(max_dim)]
        out["stats"] = {k: 0 for k in STATS}
        return out
    keep = []
    if edges is not None:
        ei, ej, ed = (np.ascontiguousarray(np.asarray(a, dtype=t))
                      for a, t in zip(edges, (np.int32, np.int32, np.float32)))
        if not (ei.shape == ej.shape == ed.shape) or ei.ndim != 1:
            raise ValueError("edges must be three 1-D arrays of one length")
        if ei.size and not (ei < ej).all():
            raise ValueError("edges must have edge_i < edge_j")
        keep = [ei, ej, ed]
        eargs = (ei.ctypes.data, ej.ctypes.data, ed.ctypes.data, ei.size)
    else:
        eargs = (None, None, None, -1)
    h = lib.orc_run_ex(pts.ctypes.data, n, d, float(threshold_sq_f32(max_edge_length)),
                       int(max_dim), 1 if exact else 0, *eargs, int(max_col), int(max_pool))
    del keep
    try:
        for idx, name in enumerate(NAMES):
            ln = lib.orc_len(h, idx)
            dt = _DTYPES.get(name, np.int32)
            shape = (ln, 2) if name in ("h0", "h1", "h2") else (ln,)
            arr = np.empty(shape, dtype=dt)
            if arr.size:
                lib.orc_get(h, idx, arr.ctypes.data)
            out[name] = arr
    finally:
        lib.orc_free(h)
    out["stats"] = dict(zip(STATS, (int(x) for x in out["stats"])))
    dg = [out["h0"]]
    if max_dim >= 1:
        dg.append(out["h1"])
    if max_dim >= 2:
        dg.append(out["h2"])
    out["diagrams"] = dg
    return out


def dist_sq(points, i: int, j: int) -> np.float32:
    lib = _load()
    pts = np.ascontiguousarray(np.asarray(points, dtype=np.float32))
    return np.float32(lib.orc_dist_sq(pts.ctypes.data, pts.shape[1], i, j))


def vertex_neighbours(points, max_edge_length: float, v: int):
    """(neighbour ids ascending, their FMA-chain dsq) of vertex v: every u != v
    with dsq < fp32(t^2).  O(n d) per call."""
    lib = _load()
    pts = np.ascontiguousarray(np.asarray(points, dtype=np.float32))
    n, d = pts.shape
    thr = float(threshold_sq_f32(max_edge_length))
    m = lib.orc_vertex_neighbours(pts.ctypes.data, n, d, thr, int(v), None, None, 0)
    idx = np.empty(max(m, 1), dtype=np.int32)
    dsq = np.empty(max(m, 1), dtype=np.float32)
    lib.orc_vertex_neighbours(pts.ctypes.data, n, d, thr, int(v), idx.ctypes.data,
                              dsq.ctypes.data, m)
    return idx[:m], dsq[:m]


def fmaf(a, b, c) -> np.float32:
    return np.float32(_load().orc_fmaf(float(a), float(b), float(c)))


def knn_kth_dsq(points, k: int) -> np.ndarray:
    """k-th nearest-neighbour squared distance per point (FMA-chain form)."""
    pts = np.ascontiguousarray(np.asarray(points, dtype=np.float32))
    n, d = pts.shape
    out = np.empty(n, dtype=np.float32)
    if _load().orc_knn_kth(pts.ctypes.data, n, d, int(k), out.ctypes.data) != 0:
        raise ValueError("knn_kth_dsq: need 1 <= k <= n-1")
    return out


def sqrt_f32(x) -> np.ndarray:
    """Correctly rounded FP32 square root (sqrt in double, then rounded)."""
    return np.sqrt(np.asarray(x, dtype=np.float32).astype(np.float64)).astype(np.float32)


def quantile_f32(values, q: float) -> np.float32:
    """torch.quantile(values_f32, q), linear interpolation, restated in FP32
    (ATen quantile: rank = q * (n - 1) in the input dtype, lerp between the
    order statistics below and above, with torch's lerp:
    a + w (b - a) for |w| < 0.5, else b - (b - a)(1 - w))."""
    v = np.sort(np.asarray(values, dtype=np.float32))
    n = v.size
    qf = np.float32(q)
    rank = np.float32(qf * np.float32(n - 1))
    lo = int(np.floor(rank))
    hi = int(np.ceil(rank))
    w = np.float32(rank - np.float32(lo))
    a, b = v[lo], v[hi]
    if abs(w) < np.float32(0.5):
        return np.float32(a + np.float32(w * np.float32(b - a)))
    return np.float32(b - np.float32(np.float32(b - a) * np.float32(np.float32(1.0) - w)))


def auto_threshold(points, k: int = 20, percentile: float = 95.0) -> float:
    """Restated auto_threshold (threshold.py:20-78) on the FMA-form distances:
    k clamped to [1, n-1], percentile to [0, 100]; 0.0 for n <= 1."""
    pts = np.asarray(points, dtype=np.float32)
    n = pts.shape[0]
    if n <= 1:
        return 0.0
    k = max(1, min(int(k), n - 1))
    p = max(0.0, min(100.0, float(percentile)))
    kth = sqrt_f32(knn_kth_dsq(pts, k))
    return float(quantile_f32(kth, p / 100.0))


def enclosing_radius(points) -> float:
    """Restated enclosing_radius (threshold.py:81-116, 179-200) on the
    FMA-form distances: sqrt(min_x max_y dsq); 0.0 for n <= 1."""
    pts = np.ascontiguousarray(np.asarray(points, dtype=np.float32))
    n, d = pts.shape
    if n <= 1:
        return 0.0
    far = np.empty(n, dtype=np.float32)
    _load().orc_far_dsq(pts.ctypes.data, n, d, far.ctypes.data)
    return float(sqrt_f32(far.min()))
