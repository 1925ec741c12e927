"""hipGraph-captured H0 pipeline for repeated clouds of one shape.

``rips_persistence(points, t, max_dim=0)`` launches about 70 kernels per call
and reads the edge count back to size the sort and CSR stages.  For a loop
over many clouds of the same (n, d) and threshold (a training loop, a
stream of frames), ``RipsH0Graph`` captures the whole H0 chain once -- screen,
sort, vertex CSR, Boruvka, H0 diagram -- into one hipGraph
(fph_h0_graph_create) sized by an edge capacity, and replays it for every
new cloud: one graph launch and one count read per call.  When a cloud has
more edges than the capacity, the graph is re-captured with room for them
and replayed, so the result always equals the eager path's.  Without an
explicit capacity, the first call screens the cloud once to count its
edges and sizes the slots at 1.0625x that count (+256): every stage of the
graph then runs over about as many slots as the eager path has edges.

Not in the reference, which drives every stage from Python on each call
(persistence.py:140-186).
"""
from __future__ import annotations

import ctypes

import torch

from flash_ph import _lib, _pipeline

MAX_N = 92_682


class RipsH0Graph:
    """H0 diagrams of (n, d) clouds at one threshold through a replayed
    hipGraph.  ``g(points)`` returns the (n, 2) float32 H0 diagram, equal to
    ``rips_persistence(points, max_edge_length, max_dim=0)[0]``."""

    def __init__(self, n: int, d: int, max_edge_length: float, device=None,
                 edge_capacity: int | None = None):
        if max_edge_length < 0:
            raise ValueError(f"max_edge_length must be non-negative, got {max_edge_length}")
        n, d = int(n), int(d)
        if not 2 <= n <= MAX_N:
            raise ValueError(f"RipsH0Graph needs 2 <= n <= {MAX_N}, got n={n}")
        if d < 1:
            raise ValueError(f"RipsH0Graph needs d >= 1, got d={d}")
        _lib.require_gpu()
        self.n, self.d = n, d
        self.max_edge_length = float(max_edge_length)
        self.thr_sq = _pipeline.threshold_sq_f32(max_edge_length)
        self.screen = 1 if _pipeline.use_mfma_screen(n, d) else 0
        self.device = torch.device(device) if device is not None else torch.device("cuda")
        self.points = torch.zeros(n, d, dtype=torch.float32, device=self.device)
        self.cap = int(edge_capacity) if edge_capacity else 0
        self.captures = 0
        self.E = 0
        self._handle = None

    # -- capture ----------------------------------------------------------
    def _capture(self):
        self.close()
        lib = _lib.load()
        dev, n, cap = self.device, self.n, self.cap
        e = lambda dt, *sh: torch.empty(*sh, dtype=dt, device=dev)
        # [0] edge count, [1] screen candidates, [2] low half: convergence
        # flag -- read back together, one host copy per replay
        self.state = torch.zeros(3, dtype=torch.int64, device=dev)
        self.count = self.state[0:2]
        self.edge_i, self.edge_j = e(torch.int32, cap), e(torch.int32, cap)
        self.edge_dsq, self.edge_filt = e(torch.float32, cap), e(torch.float32, cap)
        self.adj_ptr = e(torch.int32, n + 1)
        self.adj_idx, self.adj_rank = e(torch.int32, 2 * cap), e(torch.int32, 2 * cap)
        self.adj_dsq = e(torch.float32, 2 * cap)
        self.mst_mask = e(torch.uint8, cap)
        self.h0 = e(torch.float32, n, 2)
        p = _lib.ptr
        bufs = (p(self.count), p(self.edge_i), p(self.edge_j), p(self.edge_dsq),
                p(self.edge_filt), p(self.adj_ptr), p(self.adj_idx), p(self.adj_rank),
                p(self.adj_dsq), p(self.mst_mask), p(self.h0), p(self.state) + 16)
        head = (p(self.points), n, self.d, self.thr_sq, self.screen, cap)
        nb = ctypes.c_size_t(0)
        _lib.check(lib.fph_h0_pipeline(*head, *bufs, None, ctypes.byref(nb), None),
                   "fph_h0_pipeline")
        self.ws = e(torch.uint8, max(int(nb.value), 1))
        # the capture only records; finish pending writes to these buffers first
        torch.cuda.synchronize(dev)
        h = ctypes.c_void_p()
        _lib.check(lib.fph_h0_graph_create(*head, *bufs, p(self.ws), int(nb.value),
                                           ctypes.addressof(h)), "fph_h0_graph_create")
        self._handle = h
        self.captures += 1

    @staticmethod
    def _headroom(E: int) -> int:
        return E + E // 16 + 256

    def _count_edges(self) -> int:
        """Edge count of the current points (one eager screen)."""
        cx = _pipeline.build_edges(self.points, self.max_edge_length)
        return cx.E

    def _launch(self):
        _lib.check(_lib.load().fph_graph_launch(self._handle, _lib.stream_of(self.device)),
                   "fph_graph_launch")

    # -- use --------------------------------------------------------------
    def run(self, points) -> torch.Tensor:
        x = torch.as_tensor(points)
        if tuple(x.shape) != (self.n, self.d):
            raise ValueError(f"expected points of shape ({self.n}, {self.d}), "
                             f"got {tuple(x.shape)}")
        self.points.copy_(x.to(self.device).float())
        if self._handle is None:
            if self.cap <= 0:
                self.cap = self._headroom(self._count_edges())
            self._capture()
        self._launch()
        E, _, conv = self.state.tolist()
        if E > self.cap:
            # more edges than slots: re-capture with room and replay
            self.cap = self._headroom(E)
            self._capture()
            self._launch()
            E, _, conv = self.state.tolist()
            if E > self.cap:
                raise _lib.FlashPHError("RipsH0Graph: edge count grew on replay")
        if not conv & 0xFFFFFFFF:
            raise _lib.FlashPHError("Boruvka did not converge")
        self.E = E
        return self.h0.clone()

    __call__ = run

    def edges(self):
        """(edge_i, edge_j, edge_dsq) of the last cloud, in filtration order."""
        E = self.E
        return self.edge_i[:E], self.edge_j[:E], self.edge_dsq[:E]

    def close(self):
        if self._handle is not None:
            torch.cuda.synchronize(self.device)
            _lib.load().fph_graph_destroy(self._handle)
            self._handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
