"""Loader of libflashph_cpu.so, the host (CPU) pipeline used for point
clouds held in host memory, as the reference computes CPU tensors on the
CPU (rips.py:221-233, persistence.py:197-206).  C++ with OpenMP; same
filtration, orders and pairing as the gfx950 library, so its diagrams are
bit-identical.  It never stands in for the GPU library: CUDA tensors always
run on the GPU, and a missing GPU library raises there.
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "_lib", "libflashph_cpu.so")

FPHC_OK = 0
FPHC_ERR_LIMIT = -3
FPHC_ERR_MEM = -5
STATS = ["E", "T", "Q", "xor1", "xor2", "drops", "truncations"]

_lib = None


def load():
    """Load and bind libflashph_cpu.so; raises if it is not built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise ImportError(
            f"flash_ph: the CPU library {LIB_PATH} is not built; run "
            "`python flash-ph_amd/build.py`")
    lib = ctypes.CDLL(LIB_PATH)
    P, I64, I32, F32 = ctypes.c_void_p, ctypes.c_int64, ctypes.c_int, ctypes.c_float
    lib.fphc_version.restype = ctypes.c_char_p
    lib.fphc_version.argtypes = []
    lib.fphc_rips.restype = I32
    lib.fphc_rips.argtypes = [P, I64, I32, F32, I32, I32, I32, I64, I64, I32,
                              ctypes.POINTER(P)]
    lib.fphc_rows.restype = I64
    lib.fphc_rows.argtypes = [P, I32]
    lib.fphc_copy.restype = I32
    lib.fphc_copy.argtypes = [P, I32, P]
    lib.fphc_stats.restype = I32
    lib.fphc_stats.argtypes = [P, P]
    lib.fphc_free.restype = None
    lib.fphc_free.argtypes = [P]
    lib.fphc_filtration.restype = I32
    lib.fphc_filtration.argtypes = [P, I64, I32, F32, I32, I32, ctypes.POINTER(P)]
    lib.fphc_filtration_edges.restype = I64
    lib.fphc_filtration_edges.argtypes = [P]
    lib.fphc_filtration_copy.restype = I32
    lib.fphc_filtration_copy.argtypes = [P, P, P, P, P, P, P, P]
    lib.fphc_filtration_free.restype = None
    lib.fphc_filtration_free.argtypes = [P]
    lib.fphc_knn_kth.restype = I32
    lib.fphc_knn_kth.argtypes = [P, I64, I32, I32, P, I32]
    lib.fphc_far_dsq.restype = I32
    lib.fphc_far_dsq.argtypes = [P, I64, I32, P, I32]
    lib.fphc_apparent_scan.restype = I32
    lib.fphc_apparent_scan.argtypes = [P, P, P, I64, P, P, P, P, P, I32]
    lib.fphc_apparent_scan_h2.restype = I32
    lib.fphc_apparent_scan_h2.argtypes = [P, P, P, P, I64, P, P, P, P, P, I32]
    _lib = lib
    return lib


def _check(rc: int, what: str):
    if rc == FPHC_OK:
        return
    if rc == FPHC_ERR_LIMIT:
        raise ValueError(f"{what}: simplex count exceeds the int32 rank range")
    if rc == FPHC_ERR_MEM:
        raise MemoryError(f"{what}: host memory exhausted")
    raise RuntimeError(f"{what} failed with status {rc}")


def rips(points: np.ndarray, thr_sq: float, max_dim: int, exact_h1: bool = False,
         exact_colex: bool = False, max_col: int = 0, max_pool: int = 0,
         threads: int = 0, want_stats: bool = False):
    """Diagrams [H0, ..., H_max_dim] as (K, 2) float32 arrays."""
    lib = load()
    pts = np.ascontiguousarray(points, dtype=np.float32)
    n, d = pts.shape
    h = ctypes.c_void_p()
    _check(lib.fphc_rips(pts.ctypes.data, n, max(d, 1), float(thr_sq), int(max_dim),
                         int(bool(exact_h1)), int(bool(exact_colex)), int(max_col),
                         int(max_pool), int(threads), ctypes.byref(h)), "fphc_rips")
    try:
        out = []
        for k in range(max_dim + 1):
            m = lib.fphc_rows(h, k)
            a = np.empty((m, 2), dtype=np.float32)
            if m:
                _check(lib.fphc_copy(h, k, a.ctypes.data), "fphc_copy")
            out.append(a)
        if want_stats:
            st = np.zeros(8, dtype=np.int64)
            lib.fphc_stats(h, st.ctypes.data)
            return out, dict(zip(STATS, st.tolist()))
        return out
    finally:
        lib.fphc_free(h)


def knn_kth_dsq(points: np.ndarray, k: int, threads: int = 0) -> np.ndarray:
    lib = load()
    pts = np.ascontiguousarray(points, dtype=np.float32)
    n, d = pts.shape
    out = np.empty(n, dtype=np.float32)
    _check(lib.fphc_knn_kth(pts.ctypes.data, n, d, int(k), out.ctypes.data, int(threads)),
           "fphc_knn_kth")
    return out


def far_dsq(points: np.ndarray, threads: int = 0) -> np.ndarray:
    lib = load()
    pts = np.ascontiguousarray(points, dtype=np.float32)
    n, d = pts.shape
    out = np.empty(n, dtype=np.float32)
    _check(lib.fphc_far_dsq(pts.ctypes.data, n, d, out.ctypes.data, int(threads)),
           "fphc_far_dsq")
    return out


def filtration(points: np.ndarray, thr_sq: float, exact_colex: bool = False,
               threads: int = 0) -> dict:
    """Edges in filtration order and the vertex CSR as numpy arrays:
    edge_i, edge_j, edge_dsq, adj_ptr, adj_idx, adj_rank, adj_dsq."""
    lib = load()
    pts = np.ascontiguousarray(points, dtype=np.float32)
    n, d = pts.shape
    h = ctypes.c_void_p()
    _check(lib.fphc_filtration(pts.ctypes.data, n, max(d, 1), float(thr_sq),
                               int(bool(exact_colex)), int(threads), ctypes.byref(h)),
           "fphc_filtration")
    try:
        E = lib.fphc_filtration_edges(h)
        out = {"edge_i": np.empty(E, np.int32), "edge_j": np.empty(E, np.int32),
               "edge_dsq": np.empty(E, np.float32), "adj_ptr": np.empty(n + 1, np.int32),
               "adj_idx": np.empty(2 * E, np.int32), "adj_rank": np.empty(2 * E, np.int32),
               "adj_dsq": np.empty(2 * E, np.float32)}
        order = ["edge_i", "edge_j", "edge_dsq", "adj_ptr", "adj_idx", "adj_rank", "adj_dsq"]
        _check(lib.fphc_filtration_copy(h, *(out[k].ctypes.data for k in order)),
               "fphc_filtration_copy")
        return out
    finally:
        lib.fphc_filtration_free(h)


def _i32(a) -> np.ndarray:
    return np.ascontiguousarray(a, dtype=np.int32)


def apparent_scan(edge_i, edge_j, nonmst_ranks, adj_ptr, adj_idx, adj_rank,
                  threads: int = 0) -> tuple[np.ndarray, np.ndarray]:
    """H1 apparent-pair scan, complete (fphc_apparent_scan)."""
    lib = load()
    a = [_i32(x) for x in (edge_i, edge_j, nonmst_ranks, adj_ptr, adj_idx, adj_rank)]
    K = a[2].shape[0]
    flag = np.zeros(K, dtype=np.int32)
    colex = np.zeros(K, dtype=np.int64)
    _check(lib.fphc_apparent_scan(a[0].ctypes.data, a[1].ctypes.data, a[2].ctypes.data, K,
                                  a[3].ctypes.data, a[4].ctypes.data, a[5].ctypes.data,
                                  flag.ctypes.data, colex.ctypes.data, int(threads)),
           "fphc_apparent_scan")
    return flag, colex


def apparent_scan_h2(tri_v0, tri_v1, tri_v2, tri_max_rank, adj_ptr, adj_idx, adj_rank,
                     threads: int = 0) -> tuple[np.ndarray, np.ndarray]:
    """H2 apparent-pair scan, complete (fphc_apparent_scan_h2)."""
    lib = load()
    a = [_i32(x) for x in (tri_v0, tri_v1, tri_v2, tri_max_rank, adj_ptr, adj_idx, adj_rank)]
    T = a[0].shape[0]
    flag = np.zeros(T, dtype=np.int32)
    colex = np.zeros(T, dtype=np.int64)
    _check(lib.fphc_apparent_scan_h2(a[0].ctypes.data, a[1].ctypes.data, a[2].ctypes.data,
                                     a[3].ctypes.data, T, a[4].ctypes.data, a[5].ctypes.data,
                                     a[6].ctypes.data, flag.ctypes.data, colex.ctypes.data,
                                     int(threads)), "fphc_apparent_scan_h2")
    return flag, colex
