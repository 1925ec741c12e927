"""Loader for libflashph.so, the gfx950 HIP library behind flash_ph.

The library is built in-tree (flash-ph_amd/build.py) and bound with ctypes:
every entry point takes device pointers, sizes and a hipStream_t, and returns
an int status.  There is no CPU fallback: when the library or a GPU is
missing, the calls below raise.
"""
from __future__ import annotations

import ctypes
import os

import torch

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "_lib", "libflashph.so")

FPH_OK = 0
FPH_ERR_ARG = -1
FPH_ERR_CAPACITY = -2
FPH_ERR_LIMIT = -3
FPH_ERR_HIP = -4

_P = ctypes.c_void_p
_I64 = ctypes.c_int64
_I32 = ctypes.c_int
_F32 = ctypes.c_float
_F64 = ctypes.c_double
_SZ = ctypes.c_size_t
_PSZ = ctypes.POINTER(ctypes.c_size_t)

# name -> (restype, argtypes); must match include/flashph.h
SIGNATURES = {
    "fph_version": (ctypes.c_char_p, []),
    "fph_edge_screen": (_I32, [_P, _I64, _I32, _F32, _P, _P, _I64, _P, _P]),
    "fph_edge_screen_mfma": (_I32, [_P, _I64, _I32, _F32, _P, _P, _I64, _P, _P, _P, _PSZ, _P]),
    "fph_edge_screen_mfma_band": (_I32, [_P, _I64, _I32, _F32, _P, _P, _I64, _P, _P, _P, _I64,
                                         _P, _P, _PSZ, _P]),
    "fph_pairs_fp64_check": (_I32, [_P, _I64, _I32, _F64, _F32, _P, _I64, _P, _P]),
    "fph_edge_sort": (_I32, [_P, _P, _I64, _I64, _P, _P, _P, _P, _P, _P, _P, _PSZ, _P]),
    "fph_edge_sort_colex64": (_I32, [_P, _P, _I64, _I64, _P, _P, _P, _P, _P, _P, _P, _PSZ, _P]),
    "fph_vertex_csr": (_I32, [_P, _P, _P, _I64, _I64, _P, _P, _P, _P, _P, _PSZ, _P]),
    "fph_coo_export": (_I32, [_P, _P, _P, _I64, _I64, _P, _P, _P, _P]),
    "fph_boruvka_rounds": (_I32, [_I64]),
    "fph_boruvka_scratch_bytes": (_SZ, [_I64]),
    "fph_boruvka_mst": (_I32, [_P, _P, _I64, _I64, _P, _P, _P, _P, _SZ, _P]),
    "fph_boruvka_mst_noalloc": (_I32, [_P, _P, _I64, _I64, _P, _P, _P, _P]),
    "fph_h0_from_keys": (_I32, [_P, _I64, _I64, _P, _P, _P, _PSZ, _P]),
    "fph_h0_diagram": (_I32, [_P, _P, _I64, _I64, _P, _P, _P, _P, _PSZ, _P]),
    "fph_triangle_count": (_I32, [_P, _P, _I64, _P, _P, _P, _P, _P, _P, _P, _P, _PSZ, _P]),
    "fph_triangle_emit": (_I32, [_P, _P, _I64, _P, _P, _P, _P, _P, _P, _P, _P, _P, _P, _P]),
    "fph_cofacets_h1": (_I32, [_P, _P, _I64, _I64, _P, _P, _P, _P, _P, _P, _P, _I64, _P, _P,
                               _PSZ, _P]),
    "fph_apparent": (_I32, [_I64, _I64, _P, _P, _P, _P, _I64, _I64, _P, _P, _P, _P, _P, _P]),
    "fph_apparent_scan": (_I32, [_P, _P, _P, _I64, _P, _P, _P, _I64, _P, _P, _P]),
    "fph_apparent_scan_h2": (_I32, [_P, _P, _P, _P, _I64, _P, _P, _P, _I64, _P, _P, _P]),
    "fph_reduce_scratch_bytes": (_SZ, [_I64]),
    "fph_reduce": (_I32, [_I64, _P, _P, _I64, _I64, _P, _I64, _P, _P, _P, _P, _P, _P, _I64, _P,
                          _SZ, _P, _I32, _P]),
    "fph_reduce_replay": (_I32, [_I64, _P, _P, _I64, _I64, _P, _P, _P, _P, _P, _P, _I64, _P,
                                 _SZ, _P, _I32, _P]),
    "fph_reduce_serial": (_I32, [_I64, _P, _P, _I64, _I64, _P, _I64, _P, _I64, _P, _P, _P, _P,
                                 _P, _P, _I64, _P, _SZ, _P, _I32, _I64, _P]),
    "fph_reduce_pipe": (_I32, [_I64, _P, _P, _I64, _I64, _P, _I64, _P, _I64, _I64, _P, _P, _P,
                               _P, _P, _P, _I64, _P, _SZ, _P, _PSZ, _P, _I32, _I64, _I32, _P]),
    "fph_diagram": (_I32, [_I64, _P, _P, _P, _P, _P, _I64, _P, _P, _I64, _I32, _P, _P, _P, _I64,
                           _P, _P, _P, _P, _PSZ, _P]),
    "fph_clear_pivots": (_I32, [_I64, _P, _P, _I64, _I64, _P, _P]),
    "fph_tet_count": (_I32, [_I64, _P, _P, _P, _P, _P, _P, _P, _P, _PSZ, _P]),
    "fph_tet_emit": (_I32, [_I64, _P, _P, _P, _P, _P, _P, _P, _P, _P, _P, _P, _P, _P, _P,
                            _P]),
    "fph_cofacets_h2_count": (_I32, [_I64, _P, _P, _P, _P, _P, _P, _P, _P, _P, _P, _PSZ,
                                     _P]),
    "fph_cofacets_h2": (_I32, [_I64, _I64, _I64, _P, _P, _P, _P, _P, _P, _P, _P, _P, _P, _P,
                               _I64, _P, _P, _PSZ, _P]),
    "fph_knn_threshold_batch": (_I32, [_P, _P, _P, _P, _P, _I64, _I64, _I64, _I32, _F64, _I32,
                                       _P, _P, _P, _P, _PSZ, _P]),
    "fph_enclosing_radius_batch": (_I32, [_P, _P, _P, _P, _P, _I64, _I64, _I32, _P, _P, _PSZ,
                                          _P]),
    "fph_backward_scatter": (_I32, [_P, _P, _P, _P, _I64, _I64, _P, _I64, _P, _P, _P, _P, _P,
                                    _I64, _P, _P, _PSZ, _P]),
    "fph_edge_screen_batch": (_I32, [_P, _P, _P, _P, _P, _P, _I64, _I64, _P, _P, _I64, _P, _P]),
    "fph_edge_sort_batch": (_I32, [_P, _P, _I64, _P, _I64, _P, _P, _P, _P, _P, _P, _PSZ, _P]),
    "fph_h0_diagram_batch": (_I32, [_P, _P, _I64, _P, _P, _I64, _I64, _P, _P, _PSZ, _P]),
    "fph_h0_pipeline": (_I32, [_P, _I64, _I32, _F32, _I32, _I64, _P, _P, _P, _P, _P, _P, _P,
                               _P, _P, _P, _P, _P, _P, _PSZ, _P]),
    "fph_h0_graph_create": (_I32, [_P, _I64, _I32, _F32, _I32, _I64, _P, _P, _P, _P, _P, _P,
                                   _P, _P, _P, _P, _P, _P, _P, _SZ, _P]),
    "fph_graph_launch": (_I32, [_P, _P]),
    "fph_graph_destroy": (_I32, [_P]),
    "fph_autotune_candidates": (_I32, [_I32, _P, _I32]),
    "fph_autotune_set": (_I32, [_I32, _I64, _I64, _I32]),
    "fph_autotune_query": (_I32, [_I32, _I64, _I64]),
    "fph_autotune_reset": (None, []),
}

_lib = None


class FlashPHError(RuntimeError):
    """A libflashph entry point returned an error status."""


def available() -> bool:
    return os.path.exists(LIB_PATH)


def load():
    """Load and bind libflashph.so; raises if it is not built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise ImportError(
            f"flash_ph: the HIP library {LIB_PATH} is not built; run "
            "`python flash-ph_amd/build.py` (hipcc, gfx950)")
    lib = ctypes.CDLL(LIB_PATH)
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


def require_gpu() -> None:
    if not torch.cuda.is_available():
        raise RuntimeError(
            "flash_ph runs its Vietoris-Rips pipeline on a ROCm GPU (HIP, gfx950); "
            "no GPU is visible to torch")
    load()


def check(status: int, what: str) -> None:
    if status == FPH_OK:
        return
    if status == FPH_ERR_LIMIT:
        raise ValueError(f"{what}: size exceeds the int32 index range")
    names = {FPH_ERR_ARG: "invalid argument", FPH_ERR_CAPACITY: "capacity exceeded",
             FPH_ERR_HIP: "HIP runtime error"}
    raise FlashPHError(f"{what} failed: {names.get(status, status)}")


def ptr(t) -> int | None:
    """Device address of a tensor (None for absent optional arguments)."""
    if t is None:
        return None
    return t.data_ptr()


def stream_of(device) -> int:
    return torch.cuda.current_stream(device).cuda_stream


class Workspace:
    """One growing device scratch buffer per device, reused by consecutive
    stream-ordered calls on the current stream."""

    def __init__(self):
        self._bufs: dict = {}

    def get(self, nbytes: int, device) -> torch.Tensor:
        key = (device.type, device.index)
        buf = self._bufs.get(key)
        if buf is None or buf.numel() < nbytes:
            size = max(int(nbytes), 1 << 20)
            if buf is not None:
                size = max(size, 2 * buf.numel())
            buf = torch.empty(size, dtype=torch.uint8, device=device)
            self._bufs[key] = buf
        return buf


WORKSPACE = Workspace()


def call_ws(name: str, args_before, args_after, device):
    """Call an entry point that takes (..., ws, ws_bytes, stream): query the
    workspace size, provide the shared buffer, and check the status."""
    lib = load()
    fn = getattr(lib, name)
    nbytes = ctypes.c_size_t(0)
    check(fn(*args_before, None, ctypes.byref(nbytes), *args_after), name)
    buf = WORKSPACE.get(nbytes.value, device)
    nbytes = ctypes.c_size_t(buf.numel())
    check(fn(*args_before, buf.data_ptr(), ctypes.byref(nbytes), *args_after), name)
