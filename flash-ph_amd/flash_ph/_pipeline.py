"""Host orchestration of the exact Vietoris-Rips pipeline on one GPU.

Mirrors the reference's control flow (persistence.py:68-206,
rips.py:98-218, reduce_h1_gpu.py:438-604) over the libflashph entry points.
All arrays stay on the device; the host reads back only the sizes that
shape the next allocation (edge count; triangle/cofacet counts; the row
counts of the H1 and H2 diagrams; tetrahedron/cofacet counts).
"""
from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np
import torch

from flash_ph import _lib
from flash_ph._lib import call_ws, check, load, ptr, stream_of

# Safety limits (kernels/triangle_kernel.py:295,377-381).
MAX_TRIANGLES = 10_000_000
# Screened-pair buffer size for the first attempt: at least this many edges,
# or 64 per point; a larger count triggers one exact-size retry.
_EDGE_CAP_MIN = 1 << 20
_EDGE_CAP_PER_POINT = 64
# Reduction pool (32-bit words): max(this, 8 x cofacet entries).
_POOL_MIN_WORDS = 1 << 20
# Seconds without progress after which a waiting wavefront of the
# reduction's second phase gives up (status 5).
_WAIT_SECONDS = 120
# Entry cap of the reference's reduction store (_numba_utils.py:11).
REF_MAX_POOL = 100_000_000
# Rows of a working column beyond which the reference truncates it
# (_numba_utils.py:10).  Both limits are module attributes so that tests can
# lower them and reach the reference's store-drop and truncation branches
# on small clouds; the reducers take them as arguments.
REF_MAX_COL = 1_048_576
# Largest reduction pool the retry loop may request (32-bit words, 64 GiB):
# past it a reduction fails with FlashPHError instead of asking for more.
MAX_POOL_WORDS = 1 << 34
# Reference semantics, one cloud: run the parallel reduction and replay the
# reference's additions from its final owners (fph_reduce_replay); its pairs
# are used when the reference's store would hold every column, else the
# serial reducer runs.  Needs the rows to fit a wavefront's bitmap.
REPLAY = True
_REPLAY_MAX_ROWS = 2496 * 32
# A replayed column waits only for earlier columns being rebuilt (at most
# milliseconds); past this bound the replay gives up and the serial reducer runs.
_REPLAY_WAIT_SECONDS = 10
# Largest replay pool (32-bit words); bigger complexes use the serial reducer.
_REPLAY_MAX_WORDS = 1 << 28


def _replay_words(cap: int, K: int) -> int:
    """Replay pool words: rebuilt columns up to the store capacity (beyond
    it the store overflows anyway), each rounded to 32-word lines, plus the
    wavefronts' partly used allocation chunks."""
    return cap + 32 * K + (1 << 18)
# Dynamic LDS words of the serial reducer (working-column bitmap + summary).
_BITMAP_LDS_WORDS = 32768
# The exact reduction of a segment with more rows than this runs in serial
# order (bitmap working column), else in parallel (fph_reduce).  The
# parallel scheme wins on sparse complexes with short columns (n=10K, d=10:
# 4 ms vs ~30 ms), the serial one on dense ones (n=4918, d=50: seconds vs a
# minute).
EXACT_SERIAL_MIN_ROWS = 100_000
# Working column of the serial reducer: 0 = bitmap (list form only where a
# column would pass the reference's MAX_COL), 1 = sorted list.
SERIAL_VARIANT = 0
# A global bitmap of the serial reducer keeps a window of rows in LDS (False:
# every flip and pivot search in global memory).  Tests: SERIAL_WINDOW_UNITS
# caps the window at that many 1024-word units (0: as many as LDS holds) so
# that small clouds move it.
SERIAL_WINDOW = True
SERIAL_WINDOW_UNITS = 0


def _serial_variant(serial_exact: bool) -> int:
    return (SERIAL_VARIANT | (2 if serial_exact else 0) | (0 if SERIAL_WINDOW else 8)
            | (min(max(int(SERIAL_WINDOW_UNITS), 0), 255) << 8))
# The reference-semantics reduction with many columns in flight
# (fph_reduce_pipe); False: one workgroup per segment (fph_reduce_serial).
# The sorted-list form (a column past MAX_COL) is always the serial one.
PIPE = True
# Tests: the pipelined reducer's working bitmaps in global memory even where
# LDS holds them.
PIPE_FORCE_GLOBAL = False
# Rows whose bitmap and summary fit the pipelined reducer's LDS arena
# (20224 words).  Larger segments keep a window of their bitmap in LDS over
# a global slot (PIPE_WINDOW; False: the whole bitmap in the slot, which
# loses to the one-workgroup reducer when many segments share the call: 64
# Clifford clouds of 2.8M rows, 362 vs 132 ms).  Tests: PIPE_WINDOW_UNITS
# caps the window (1024-word units) so that small clouds move it.
_PIPE_LDS_WORDS = 20224
_PIPE_GLOBAL_MAX_SEGS = 8
PIPE_WINDOW = True
PIPE_WINDOW_UNITS = 0
# At most this many unsettled columns per segment (0: a quarter of an XCD's
# wavefronts for bitmaps in LDS, an eighth else; 1..255 for tuning).
PIPE_INFLIGHT = 0


def _use_pipe(max_rows: int, nseg: int) -> bool:
    w = (max_rows + 31) // 32
    s = (w + 31) // 32
    if w + s <= _PIPE_LDS_WORDS:
        return True
    if PIPE_WINDOW and (((_PIPE_LDS_WORDS - s) * 32 // 33) & ~1023) >= 1024:
        return True
    return nseg <= _PIPE_GLOBAL_MAX_SEGS


def _pipe_variant(variant: int) -> int:
    """The serial reducer's variant bits with the pipelined reducer's window
    settings in place of the serial ones (bit 3: no window, bits 8..15: the
    window cap, bits 16..23: the in-flight cap)."""
    v = variant & ~(8 | (0xFF << 8) | (0xFF << 16))
    if not PIPE_WINDOW:
        v |= 8
    v |= min(max(int(PIPE_WINDOW_UNITS), 0), 255) << 8
    v |= min(max(int(PIPE_INFLIGHT), 0), 255) << 16
    return v | (4 if PIPE_FORCE_GLOBAL else 0)
# When a list, every reduction attempt appends (pool words, status words).
REDUCE_LOG = None


class StageTimer:
    """Optional HIP-event timing of named launches on the current stream.

    Disabled by default; ``enable()`` makes the pipeline record an event pair
    around each named entry-point call (used by bench.py for live per-kernel
    timing)."""

    def __init__(self):
        self.enabled = False
        self.records: dict = {}

    def enable(self, on: bool = True):
        self.enabled = on
        self.records = {}

    def start(self, name, device):
        if not self.enabled:
            return None
        ev = torch.cuda.Event(enable_timing=True)
        ev.record(torch.cuda.current_stream(device))
        return (name, ev, device)

    def stop(self, tok):
        if tok is None:
            return
        name, ev0, device = tok
        ev1 = torch.cuda.Event(enable_timing=True)
        ev1.record(torch.cuda.current_stream(device))
        self.records.setdefault(name, []).append((ev0, ev1))

    def summary(self) -> dict:
        """name -> (launches, mean ms); synchronises."""
        torch.cuda.synchronize()
        return {k: (len(v), sum(a.elapsed_time(b) for a, b in v) / len(v))
                for k, v in self.records.items()}


TIMER = StageTimer()


# Screening kernel: "auto" picks the split-bf16 matrix-core screen for wide
# points (d >= 32) and large clouds, the direct FP32 screen otherwise; both
# give identical edges.  "direct" / "mfma" force one.
SCREEN = "auto"
_MFMA_MIN_D = 32
_MFMA_MIN_N = 4096


def use_mfma_screen(n: int, d: int) -> bool:
    if SCREEN == "mfma":
        return True
    if SCREEN == "direct":
        return False
    return d >= _MFMA_MIN_D and n >= _MFMA_MIN_N


def threshold_sq_f32(max_edge_length: float) -> float:
    """t*t in Python double rounded to FP32, as the reference's Triton scalar
    argument (rips.py:125, edge_kernel.py:125)."""
    return float(np.float32(float(max_edge_length) * float(max_edge_length)))


def _empty(dtype, device, *shape):
    return torch.empty(*shape, dtype=dtype, device=device)


@dataclass
class Complex:
    """Device state of one point cloud's filtration (all tensors on GPU)."""
    n: int
    device: torch.device
    E: int = 0
    edge_i: torch.Tensor | None = None
    edge_j: torch.Tensor | None = None
    edge_dsq: torch.Tensor | None = None
    edge_filt: torch.Tensor | None = None
    adj_ptr: torch.Tensor | None = None
    adj_idx: torch.Tensor | None = None
    adj_rank: torch.Tensor | None = None
    adj_dsq: torch.Tensor | None = None
    mst_mask: torch.Tensor | None = None
    comp_label: torch.Tensor | None = None
    mst_idx: torch.Tensor | None = None
    h0: torch.Tensor | None = None
    info: dict = field(default_factory=dict)
    # H1
    T: int = 0
    C1: int = 0
    tri_off: torch.Tensor | None = None
    tri_w: torch.Tensor | None = None
    tri_e: torch.Tensor | None = None
    tri_ru: torch.Tensor | None = None
    tri_rv: torch.Tensor | None = None
    tri_cmf: torch.Tensor | None = None
    cof_off: torch.Tensor | None = None
    cof_data: torch.Tensor | None = None
    kind1: torch.Tensor | None = None
    pivot1: torch.Tensor | None = None
    h1: torch.Tensor | None = None
    # H2
    Q: int = 0
    C2: int = 0
    cleared: torch.Tensor | None = None
    tet_off: torch.Tensor | None = None
    tet_key: torch.Tensor | None = None
    tet_e: torch.Tensor | None = None
    tet_cmf: torch.Tensor | None = None
    cof2_off: torch.Tensor | None = None
    cof2_data: torch.Tensor | None = None
    kind2: torch.Tensor | None = None
    pivot2: torch.Tensor | None = None
    h2: torch.Tensor | None = None
    # batches: per-cloud edge ranges (device, C+1) and vertex offsets (host)
    edge_off: torch.Tensor | None = None
    vbase_host: list | None = None
    # gradients: per diagram row, the birth and death edges (-1: none)
    want_edges: bool = False
    h1_edges: object = None
    h2_edges: object = None


def build_edges(pts: torch.Tensor, max_edge_length: float, colex64: bool = False) -> Complex:
    """Stage (a): screened edges in filtration order plus the vertex CSR.
    colex64: ties in dsq by the full colex index (fph_edge_sort_colex64)
    instead of the reference's truncated key and emission order; the same
    order for n <= 92,682."""
    lib = load()
    if pts.dtype != torch.float32 or not pts.is_contiguous():
        raise ValueError("build_edges needs a contiguous float32 (n, d) tensor")
    dev = pts.device
    s = stream_of(dev)
    n, d = pts.shape
    cx = Complex(n=n, device=dev)
    thr = threshold_sq_f32(max_edge_length)
    cap = max(_EDGE_CAP_MIN, _EDGE_CAP_PER_POINT * n)
    cnt = _empty(torch.int64, dev, 2)
    mfma = use_mfma_screen(n, d)
    while True:
        keys = _empty(torch.int64, dev, cap)
        vals = _empty(torch.int64, dev, cap)
        tok = TIMER.start("edge_screen", dev)
        if mfma:
            call_ws("fph_edge_screen_mfma",
                    (ptr(pts), n, d, thr, ptr(keys), ptr(vals), cap, ptr(cnt),
                     cnt.data_ptr() + 8), (s,), dev)
        else:
            check(lib.fph_edge_screen(ptr(pts), n, d, thr, ptr(keys), ptr(vals), cap, ptr(cnt),
                                      s), "fph_edge_screen")
        TIMER.stop(tok)
        E = int(cnt[0].item())
        cx.info["screen"] = "mfma" if mfma else "direct"
        cx.info["screen_candidates"] = cnt[1] if mfma else None
        if E <= cap:
            break
        cap = E
    # edge_kernel.py:203-205
    assert E < 2**31, f"Edge count {E} exceeds int32 range; reduce max_edge_length"
    cx.E = E
    cx.edge_i = _empty(torch.int32, dev, E)
    cx.edge_j = _empty(torch.int32, dev, E)
    cx.edge_dsq = _empty(torch.float32, dev, E)
    cx.edge_filt = _empty(torch.float32, dev, E)
    if E > 0:
        ktmp = _empty(torch.int64, dev, E)
        vtmp = _empty(torch.int64, dev, E)
        call_ws("fph_edge_sort_colex64" if colex64 else "fph_edge_sort",
                (ptr(keys), ptr(vals), E, n, ptr(ktmp), ptr(vtmp), ptr(cx.edge_i),
                 ptr(cx.edge_j), ptr(cx.edge_dsq), ptr(cx.edge_filt)), (s,), dev)
        del ktmp, vtmp
    del keys, vals
    cx.adj_ptr = _empty(torch.int32, dev, n + 1)
    cx.adj_idx = _empty(torch.int32, dev, 2 * E)
    cx.adj_rank = _empty(torch.int32, dev, 2 * E)
    cx.adj_dsq = _empty(torch.float32, dev, 2 * E)
    call_ws("fph_vertex_csr",
            (ptr(cx.edge_i), ptr(cx.edge_j), ptr(cx.edge_dsq), E, n, ptr(cx.adj_ptr),
             ptr(cx.adj_idx), ptr(cx.adj_rank), ptr(cx.adj_dsq)), (s,), dev)
    return cx


def fp64_screen_report(pts: torch.Tensor, max_edge_length: float) -> dict:
    """Validation mode of the edge screen: screens (the split-bf16 screen
    when use_mfma_screen picks it, listing its margin band), then re-examines
    the edges and the band in FP64 on the device (fph_pairs_fp64_check).
    Returns counts: edges, band (candidates the FP32 recompute rejected),
    edges whose exact distance is not below t ("fp32_in_exact_out"), band
    pairs whose exact distance is ("exact_in_fp32_out"), and the largest
    FP32-vs-FP64 error of the edges' dsq in ulp.  Diagnostic only: the
    filtration stays the FP32 one."""
    lib = load()
    if pts.dtype != torch.float32 or not pts.is_contiguous():
        raise ValueError("fp64_screen_report needs a contiguous float32 (n, d) tensor")
    dev = pts.device
    s = stream_of(dev)
    n, d = pts.shape
    thr = threshold_sq_f32(max_edge_length)
    mfma = use_mfma_screen(n, d)
    cap = max(_EDGE_CAP_MIN, _EDGE_CAP_PER_POINT * n)
    bcap = 1 << 16
    cnt = _empty(torch.int64, dev, 3)
    while True:
        keys = _empty(torch.int64, dev, cap)
        vals = _empty(torch.int64, dev, cap)
        band = _empty(torch.int64, dev, bcap)
        if mfma:
            call_ws("fph_edge_screen_mfma_band",
                    (ptr(pts), n, d, thr, ptr(keys), ptr(vals), cap, ptr(cnt),
                     cnt.data_ptr() + 8, ptr(band), bcap, cnt.data_ptr() + 16), (s,), dev)
            E, _, B = (int(x) for x in cnt.tolist())
        else:
            check(lib.fph_edge_screen(ptr(pts), n, d, thr, ptr(keys), ptr(vals), cap, ptr(cnt),
                                      s), "fph_edge_screen")
            E, B = int(cnt[0].item()), 0
        if E <= cap and B <= bcap:
            break
        cap, bcap = max(cap, E), max(bcap, B)
    st = _empty(torch.int64, dev, 16)
    check(lib.fph_pairs_fp64_check(ptr(pts), n, d, float(max_edge_length), thr, ptr(vals), E,
                                   ptr(st), s), "fph_pairs_fp64_check")
    check(lib.fph_pairs_fp64_check(ptr(pts), n, d, float(max_edge_length), thr, ptr(band), B,
                                   st.data_ptr() + 64, s), "fph_pairs_fp64_check")
    e, b = st[:8].tolist(), st[8:].tolist()
    return {"screen": "mfma" if mfma else "direct", "edges": E, "band": B,
            "edges_fp32_members": e[1], "band_fp32_members": b[1],
            "fp32_in_exact_out": e[3], "exact_in_fp32_out": b[4],
            "max_ulp_fp32_vs_fp64": e[5] / 1024.0}


# Largest cloud whose packed edge keys are unique and decode back to (i, j)
# (the colex field j (j - 1) / 2 + i stays below 2^32).
MAX_KEY_DECODE_N = 92682


def h0_direct(pts: torch.Tensor, max_edge_length: float) -> torch.Tensor:
    """H0 diagram straight from the screened edge keys (fph_h0_from_keys):
    Boruvka on the keys, only the forest sorted -- no edge sort, no CSR.
    Same rows as build_edges + build_mst (persistence.py:140-186), and the
    edge count may pass 2^31 (8 bytes per edge in HBM).  n <= 92,682."""
    lib = load()
    if pts.dtype != torch.float32 or not pts.is_contiguous():
        raise ValueError("h0_direct needs a contiguous float32 (n, d) tensor")
    dev = pts.device
    s = stream_of(dev)
    n, d = pts.shape
    if n > MAX_KEY_DECODE_N:
        raise ValueError(f"h0_direct needs n <= {MAX_KEY_DECODE_N}, got {n}")
    thr = threshold_sq_f32(max_edge_length)
    cap = max(_EDGE_CAP_MIN, _EDGE_CAP_PER_POINT * n)
    cnt = _empty(torch.int64, dev, 2)
    mfma = use_mfma_screen(n, d)
    while True:
        keys = _empty(torch.int64, dev, cap)
        if mfma:
            call_ws("fph_edge_screen_mfma",
                    (ptr(pts), n, d, thr, ptr(keys), None, cap, ptr(cnt), cnt.data_ptr() + 8),
                    (s,), dev)
        else:
            check(lib.fph_edge_screen(ptr(pts), n, d, thr, ptr(keys), None, cap, ptr(cnt), s),
                  "fph_edge_screen")
        E = int(cnt[0].item())
        if E <= cap:
            break
        del keys
        cap = E
    h0 = _empty(torch.float32, dev, n, 2)
    conv = _empty(torch.int32, dev, 1)
    call_ws("fph_h0_from_keys", (ptr(keys), E, n, ptr(h0), ptr(conv)), (s,), dev)
    if not int(conv.item()):
        raise _lib.FlashPHError("Boruvka did not converge")
    return h0


def build_mst(cx: Complex, want_labels: bool = False) -> None:
    """Stage (b): Boruvka forest, component labels and the H0 diagram."""
    lib = load()
    dev, n, E = cx.device, cx.n, cx.E
    s = stream_of(dev)
    cx.mst_mask = _empty(torch.uint8, dev, max(E, 1))
    cx.comp_label = _empty(torch.int64, dev, n) if want_labels else None
    conv = _empty(torch.int32, dev, 1)
    scratch = _lib.WORKSPACE.get(lib.fph_boruvka_scratch_bytes(n), dev)
    check(lib.fph_boruvka_mst(ptr(cx.edge_i), ptr(cx.edge_j), E, n, ptr(cx.mst_mask),
                              ptr(cx.comp_label), ptr(conv), ptr(scratch), scratch.numel(), s),
          "fph_boruvka_mst")
    cx.h0 = _empty(torch.float32, dev, n, 2)
    cx.mst_idx = _empty(torch.int64, dev, max(n, 1))
    nm = _empty(torch.int64, dev, 1)
    call_ws("fph_h0_diagram",
            (ptr(cx.mst_mask), ptr(cx.edge_dsq), E, n, ptr(cx.h0), ptr(cx.mst_idx), ptr(nm)),
            (s,), dev)
    cx.info["converged"] = conv
    cx.info["n_mst"] = nm


def build_h1(cx: Complex, max_triangles: int = MAX_TRIANGLES, semantics: str = "exact",
             defer: bool = False) -> None:
    """Stages (d), (c), (e) for H1 and the H1 diagram.

    semantics: "exact" (parallel reduction; the pairs of any complete
    reduction) or "reference" (the reference's capacity-limited serial
    reduction, identical to its output even where its store overflows).
    defer (one cloud): leave the reduction's status unread; build_h2 reads it
    with its own sizes (one host synchronisation instead of two)."""
    lib = load()
    dev, n, E = cx.device, cx.n, cx.E
    s = stream_of(dev)
    cx.tri_off = _empty(torch.int32, dev, E + 1)
    cx.cof_off = _empty(torch.int32, dev, E + 1)
    totals = _empty(torch.int64, dev, 2)
    tok = TIMER.start("enumerate", dev)
    call_ws("fph_triangle_count",
            (ptr(cx.edge_i), ptr(cx.edge_j), E, ptr(cx.adj_ptr), ptr(cx.adj_idx),
             ptr(cx.adj_rank), ptr(cx.mst_mask), ptr(cx.tri_off), ptr(cx.cof_off), ptr(totals)),
            (s,), dev)
    TIMER.stop(tok)
    segs = None
    # one read for the sizes and the forest's convergence flag
    conv = cx.info["converged"].long()
    if cx.edge_off is None:
        T, C, converged = (int(x) for x in torch.cat([totals, conv]).tolist())
        tmax = T
    else:
        # batch: triangles of cloud c = those whose largest edge lies in its
        # edge range; the limit applies per cloud, as for one cloud.
        eo = cx.edge_off.long()
        tc = cx.tri_off[eo[1:]] - cx.tri_off[eo[:-1]]
        host = torch.cat([totals, tc.long(), eo, conv]).tolist()
        T, C, converged = host[0], host[1], host[-1]
        nc = len(cx.vbase_host) - 1
        tcs = host[2:2 + nc]
        tmax = max(tcs) if tcs else 0
        segs = {"cols": cx.edge_off, "edges": cx.edge_off,
                "ess_asc": (tc == 0).to(torch.uint8), "cols_host": host[2 + nc:-1]}
        ecs = [b - a for a, b in zip(segs["cols_host"][:-1], segs["cols_host"][1:])]
    if not converged:
        raise _lib.FlashPHError("Boruvka did not converge")
    if tmax > max_triangles:
        raise ValueError(
            f"Triangle count {tmax} exceeds max_triangles={max_triangles}. "
            f"Reduce max_edge_length or increase max_triangles.")
    if E + T >= 2**31 - 1:
        raise ValueError("edge + triangle count exceeds the int32 rank range")
    cx.T, cx.C1 = T, C
    cx.tri_w = _empty(torch.int32, dev, max(T, 1))
    cx.tri_e = _empty(torch.int32, dev, max(T, 1))
    cx.tri_ru = _empty(torch.int32, dev, max(T, 1))
    cx.tri_rv = _empty(torch.int32, dev, max(T, 1))
    cx.tri_cmf = _empty(torch.int32, dev, max(T, 1))
    cx.cof_data = _empty(torch.int32, dev, max(C, 1))
    tok = TIMER.start("enumerate", dev)
    check(lib.fph_triangle_emit(ptr(cx.edge_i), ptr(cx.edge_j), E, ptr(cx.adj_ptr),
                                ptr(cx.adj_idx), ptr(cx.adj_rank), ptr(cx.mst_mask),
                                ptr(cx.tri_off), ptr(cx.tri_w), ptr(cx.tri_e), ptr(cx.tri_ru),
                                ptr(cx.tri_rv), ptr(cx.tri_cmf), s), "fph_triangle_emit")
    call_ws("fph_cofacets_h1",
            (ptr(cx.edge_i), ptr(cx.edge_j), E, T, ptr(cx.adj_ptr), ptr(cx.adj_idx),
             ptr(cx.adj_rank), ptr(cx.mst_mask), ptr(cx.tri_off), ptr(cx.tri_w),
             ptr(cx.cof_off), C, ptr(cx.cof_data)), (s,), dev)
    TIMER.stop(tok)
    # the reference's store counts 3 cofacet entries per triangle and all E
    # edges as columns (reduce_h1_gpu.py:523-534)
    shapes = ([(3 * T, E, T)] if segs is None else
              [(3 * t, e, t) for t, e in zip(tcs, ecs)])
    serial_exact = False
    replay_cap = None
    if semantics == "reference":
        ref_sizes = shapes
        cap = reference_pool_caps([(3 * T, E)])[0]
        if (REPLAY and segs is None and T <= _REPLAY_MAX_ROWS
                and _replay_words(cap, E) <= _REPLAY_MAX_WORDS):
            # the parallel reduction, kept when the replay shows the
            # reference's store holds every column (then its pairs are these)
            replay_cap = cap
            ref_sizes = None
    elif semantics == "exact":
        # dense complexes (long columns): the serial complete reduction
        serial_exact = max(r for _, _, r in shapes) > EXACT_SERIAL_MIN_ROWS
        ref_sizes = shapes if serial_exact else None
    else:
        raise ValueError(f"unknown reduction semantics {semantics!r}")
    args = dict(
        K=E, col_base=0, off=cx.cof_off, data=cx.cof_data, skip=cx.mst_mask, row_cmf=cx.tri_cmf,
        row_base=E, nrows=T, C=C, col_edge=None, row_edge=cx.tri_e, edge_filt=cx.edge_filt,
        E=E, ess_ascending=(T == 0), device=dev, segs=segs, want_edges=cx.want_edges)
    if replay_cap is not None:
        res = _reduce_and_diagram(**args, replay_cap=replay_cap)
        cx.info["h1_replay"] = res[3].get("replay")
        if not res[3]["replay_ok"]:   # the reference's store overflows: its own reducer
            res = _reduce_and_diagram(**args, ref_sizes=shapes)
    else:
        res = _reduce_and_diagram(**args, ref_sizes=ref_sizes, serial_exact=serial_exact,
                                  defer=defer and segs is None)
    if defer and segs is None:
        cx.kind1, cx.pivot1, info32, fin = res
        cx.info["h1_pending"] = (info32, fin)
        return
    _set_h1(cx, res, segs is not None)


def _set_h1(cx: Complex, res, batch: bool) -> None:
    cx.kind1, cx.pivot1, h1, stats = res
    cx.h1 = h1 if batch else h1[0]
    if cx.want_edges:
        re = stats.pop("row_edges")
        cx.h1_edges = re if batch else re[0]
    cx.info["h1_stats"] = stats


def reference_pool_caps(sizes) -> list:
    """Store capacity of the reference's reducer per segment, from
    (total cofacet entries, columns) (reduce_h1_gpu.py:286-289)."""
    return [min(max(2 * cof, 16 * k, 100_000), REF_MAX_POOL) for cof, k in sizes]


def _round32(x: int) -> int:
    return (x + 31) // 32 * 32


def _reduce_and_diagram(K, col_base, off, data, skip, row_cmf, row_base, nrows, C, col_edge,
                        row_edge, edge_filt, E, ess_ascending, device, segs=None,
                        ref_sizes=None, serial_exact=False, want_edges=False, defer=False,
                        replay_cap=None):
    """Apparent pairs, residual reduction and the diagram of one dimension.

    ``segs`` (batches) = dict(cols=, edges=, ess_asc= device tensors,
    cols_host= list): the columns split into independent point clouds.
    ``ref_sizes``: None for the parallel exact reduction, or per segment
    (reference cofacet entries, reference columns, rows) to run the serial
    reduction (fph_reduce_serial): the reference's capacity-limited one, or
    with ``serial_exact`` the complete one in serial order.
    ``replay_cap`` (one segment, parallel reduction): the reference's store
    capacity; the reduction is then followed by fph_reduce_replay, and
    stats["replay_ok"] says whether its pairs are the reference's.
    Returns (kind, pivot, [diagram per segment], stats)."""
    lib = load()
    s = stream_of(device)
    nseg = 1 if segs is None else len(segs["cols_host"]) - 1
    seg_cols = None if segs is None else segs["cols"]
    seg_edges = None if segs is None else segs["edges"]
    seg_ess_asc = None if segs is None else segs["ess_asc"]
    kind = _empty(torch.uint8, device, max(K, 1))
    own = _empty(torch.int64, device, max(nrows, 1))
    pivot = _empty(torch.int32, device, max(K, 1))
    res_list = _empty(torch.int32, device, max(K, 1))
    # info32: [0..8) reduction status, [8] residual count, [12..12+nseg) rows,
    # [16..24) replay status; every entry read back is written by the
    # entry points (their status resets, fph_apparent's count, fph_diagram's
    # row counts, the replay's reset), so it needs no clearing here
    info32 = _empty(torch.int32, device, max(12 + nseg, 24 if replay_cap is not None else 0))
    status_p = info32.data_ptr()
    rescnt_p = info32.data_ptr() + 32
    counts_p = info32.data_ptr() + 48
    scratch_bytes = lib.fph_reduce_scratch_bytes(K)
    if ref_sizes is None:
        # apparent columns in place (C + K words), owner lines (32 words per
        # row, used by the device-wide queue) + line-rounded reduced columns
        pool_words = C + K + 64 + 32 * nrows + max(_POOL_MIN_WORDS, 64 * K + 4 * C)
    else:
        caps = ([1 << 62] * len(ref_sizes) if serial_exact else
                reference_pool_caps([(cof, kc) for cof, kc, _ in ref_sizes]))
        base = _round32(C + K) + 32
        # fph_reduce_pipe stages every candidate from here on (over the store
        # regions and bitmaps, which only fph_reduce_serial uses)
        stage_start = base
        seg_pool, bitmaps = [], []
        row0 = row_base
        for cap, (_, kc, rows) in zip(caps, ref_sizes):
            if serial_exact:      # stored columns bump-allocated from the pool
                seg_pool += [-1, cap, row0, rows, -1]
            else:
                seg_pool += [base, cap, row0, rows, -1]
                base += _round32(cap + kc + 64)
            w = (rows + 31) // 32
            if w + (w + 31) // 32 > _BITMAP_LDS_WORDS:
                bitmaps.append((len(seg_pool) - 1, w))
            row0 += rows
        bm_lo = base
        for slot, w in bitmaps:       # global working-column bitmaps
            seg_pool[slot] = base
            base += _round32(w)
        bm_hi = base
        pool_start = base
        max_rows = max(rows for _, _, rows in ref_sizes)
        seg_pool_t = torch.tensor(seg_pool, dtype=torch.int64).to(device)
        # + two growing off-chip buffers per segment (list form only), and
        # for the complete reduction every stored column
        pool_words = pool_start + max(_POOL_MIN_WORDS, nseg * 262144)
        if serial_exact:
            pool_words += 64 * K + 4 * C
        variant = _serial_variant(serial_exact)
        if PIPE and (PIPE_FORCE_GLOBAL or _use_pipe(max_rows, nseg)):
            # fph_reduce_pipe stages every candidate, stored or not, in
            # 32-word lines from per-wavefront reservations: twice the
            # store regions (the uneven batch's heaviest cloud needs 1.5-2x)
            pool_words = min(stage_start + 2 * (pool_words - stage_start), 0xFFFFFFFF - 64)
    out = _empty(torch.float32, device, max(K, 1), 2)
    out_edges = _empty(torch.int32, device, max(K, 1), 2) if want_edges else None
    state = {"pool_words": pool_words, "variant": variant if ref_sizes is not None else 0}

    def launch():
        pool_words = state["pool_words"]
        tok = TIMER.start("apparent", device)
        check(lib.fph_apparent(K, col_base, ptr(off), ptr(data), ptr(skip), ptr(row_cmf),
                               row_base, nrows, ptr(kind), ptr(own), ptr(pivot), ptr(res_list),
                               rescnt_p, s), "fph_apparent")
        TIMER.stop(tok)
        pool = _empty(torch.int32, device, pool_words + 64)  # +64: speculative line reads
        scratch = _empty(torch.uint8, device, scratch_bytes)
        tok = TIMER.start("reduce", device)
        if ref_sizes is None:
            check(lib.fph_reduce(K, ptr(off), ptr(data), row_base, nrows, ptr(seg_cols), nseg,
                                 ptr(kind), ptr(own), ptr(pivot), ptr(res_list), rescnt_p,
                                 ptr(pool), pool_words, ptr(scratch), scratch_bytes, status_p,
                                 _WAIT_SECONDS, s), "fph_reduce")
            if replay_cap is not None:
                words = _replay_words(replay_cap, K)
                pool2 = _empty(torch.int32, device, words + 64)
                check(lib.fph_reduce_replay(K, ptr(off), ptr(data), row_base, nrows, ptr(kind),
                                            ptr(own), ptr(pivot), ptr(res_list), rescnt_p,
                                            ptr(pool2), words, ptr(scratch), scratch_bytes,
                                            info32.data_ptr() + 64, _REPLAY_WAIT_SECONDS, s),
                      "fph_reduce_replay")
        else:
            if pool_words > 0xFFFFFFFF:
                raise _lib.FlashPHError("reduction store exceeds 2^32 words")
            if (PIPE and not state["variant"] & 1
                    and (PIPE_FORCE_GLOBAL or _use_pipe(max_rows, nseg))):
                call_ws("fph_reduce_pipe",
                        (K, ptr(off), ptr(data), row_base, nrows, ptr(seg_cols), nseg,
                         ptr(seg_pool_t), max_rows, stage_start, ptr(kind), ptr(own), ptr(pivot),
                         ptr(res_list), rescnt_p, ptr(pool), pool_words, ptr(scratch),
                         scratch_bytes),
                        (status_p, _pipe_variant(state["variant"]),
                         REF_MAX_COL, _WAIT_SECONDS, s), device)
            else:
                if bm_hi > bm_lo:
                    pool[bm_lo:bm_hi].zero_()
                check(lib.fph_reduce_serial(K, ptr(off), ptr(data), row_base, nrows,
                                            ptr(seg_cols), nseg, ptr(seg_pool_t), pool_start,
                                            ptr(kind), ptr(own), ptr(pivot), ptr(res_list),
                                            rescnt_p, ptr(pool), pool_words, ptr(scratch),
                                            scratch_bytes, status_p, state["variant"],
                                            REF_MAX_COL, s), "fph_reduce_serial")
        TIMER.stop(tok)
        call_ws("fph_diagram",
                (K, ptr(kind), ptr(off), ptr(data), ptr(pivot), ptr(col_edge), row_base,
                 ptr(row_edge), ptr(edge_filt), E, int(bool(ess_ascending)), ptr(seg_cols),
                 ptr(seg_edges), ptr(seg_ess_asc), nseg, ptr(out), ptr(out_edges), counts_p),
                (s,), device)

    def finish(st):
        """Retries on a recoverable status, then the diagrams and stats."""
        while True:
            if REDUCE_LOG is not None:
                REDUCE_LOG.append((state["pool_words"], st[:9]))
            if st[0] == 0:
                break
            if st[0] == 1:
                want = max(2 * state["pool_words"], 32 * st[5] + (32 * st[5]) // 4)
                if want > MAX_POOL_WORDS:
                    raise _lib.FlashPHError(
                        f"reduction pool would need {want} words (> {MAX_POOL_WORDS}); "
                        "reduce max_edge_length")
                state["pool_words"] = want
            elif st[0] == 4 and ref_sizes is not None and (state["variant"] & 1) == 0:
                state["variant"] |= 1  # a column would pass the reference's MAX_COL: list form
            else:
                raise _lib.FlashPHError(f"fph_reduce failed with status {st[0]}")
            launch()
            st = info32.tolist()
        merged = (st[6] & 0xFFFFFFFF) | ((st[7] & 0xFFFFFFFF) << 32)
        stats = {"xors": st[2], "longest": st[3], "offchip_xors": st[4],
                 "pool_words": 32 * st[5], "pool_cap": state["pool_words"], "residual": st[8],
                 "merged": merged}
        if replay_cap is not None:
            r = st[16:24]
            held = (r[4] & 0xFFFFFFFF) | ((r[5] & 0xFFFFFFFF) << 32)
            stats["replay"] = {"status": r[0], "xors": r[2], "longest": r[3], "held": held,
                               "held_columns": r[6], "cap": replay_cap}
            stats["replay_ok"] = (r[0] == 0 and (r[1] == 1 or K == 0) and held <= replay_cap
                                  and r[3] <= REF_MAX_COL and r[6] <= K)
        counts = st[12:12 + nseg]
        starts = [0] if segs is None else segs["cols_host"][:-1]
        dgms = [out[a:a + c] for a, c in zip(starts, counts)]
        if want_edges:
            stats["row_edges"] = [out_edges[a:a + c] for a, c in zip(starts, counts)]
        return kind, pivot, dgms, stats

    launch()
    if defer:
        # the caller reads info32 together with its next sizes and calls
        # finish(that list); kind/pivot are final unless finish retries
        return kind, pivot, info32, finish
    return finish(info32.tolist())


def build_h2(cx: Complex) -> None:
    """H1 -> H2 clearing, tetrahedra, H2 columns, reduction and diagram."""
    lib = load()
    dev, E, T = cx.device, cx.E, cx.T
    s = stream_of(dev)
    cx.cleared = _empty(torch.uint8, dev, max(T, 1))
    cx.tet_off = _empty(torch.int32, dev, E + 1)
    cx.cof2_off = _empty(torch.int32, dev, T + 1)
    totals = _empty(torch.int64, dev, 2)

    def clear_and_count():
        tok = TIMER.start("apparent", dev)
        check(lib.fph_clear_pivots(E, ptr(cx.kind1), ptr(cx.pivot1), E, T, ptr(cx.cleared), s),
              "fph_clear_pivots")
        TIMER.stop(tok)
        tok = TIMER.start("enumerate", dev)
        call_ws("fph_tet_count",
                (E, ptr(cx.adj_ptr), ptr(cx.adj_idx), ptr(cx.adj_rank), ptr(cx.tri_off),
                 ptr(cx.tri_w), ptr(cx.tet_off), totals.data_ptr()), (s,), dev)
        call_ws("fph_cofacets_h2_count",
                (T, ptr(cx.edge_i), ptr(cx.edge_j), ptr(cx.adj_ptr), ptr(cx.adj_idx),
                 ptr(cx.tri_w), ptr(cx.tri_e), ptr(cx.cleared), ptr(cx.cof2_off),
                 totals.data_ptr() + 8), (s,), dev)
        TIMER.stop(tok)

    clear_and_count()
    pending = cx.info.pop("h1_pending", None)
    if pending is None:
        Q, C2 = (int(x) for x in totals.tolist())
    else:
        # the H1 status read together with these sizes; a retried H1 reduction
        # (rare: its store was too small) changes the clearing, so recount
        info32, fin = pending
        host = torch.cat([totals, info32.long()]).tolist()
        Q, C2, st1 = host[0], host[1], host[2:]
        _set_h1(cx, fin(st1), False)
        if st1[0] != 0:
            clear_and_count()
            Q, C2 = (int(x) for x in totals.tolist())
    if E + T + Q >= 2**31 - 1:
        raise ValueError("simplex count exceeds the int32 rank range")
    cx.Q, cx.C2 = Q, C2
    cx.tet_key = _empty(torch.int64, dev, max(Q, 1))
    cx.tet_e = _empty(torch.int32, dev, max(Q, 1))
    cx.tet_cmf = _empty(torch.int32, dev, max(Q, 1))
    cx.cof2_data = _empty(torch.int32, dev, max(C2, 1))
    tok = TIMER.start("enumerate", dev)
    check(lib.fph_tet_emit(E, ptr(cx.edge_i), ptr(cx.edge_j), ptr(cx.adj_ptr), ptr(cx.adj_idx),
                           ptr(cx.adj_rank), ptr(cx.tri_off), ptr(cx.tri_w), ptr(cx.tri_ru),
                           ptr(cx.tri_rv), ptr(cx.cleared), ptr(cx.tet_off), ptr(cx.tet_key),
                           ptr(cx.tet_e), ptr(cx.tet_cmf), s), "fph_tet_emit")
    call_ws("fph_cofacets_h2",
            (T, E, Q, ptr(cx.edge_i), ptr(cx.edge_j), ptr(cx.adj_ptr), ptr(cx.adj_idx),
             ptr(cx.adj_rank), ptr(cx.tri_w), ptr(cx.tri_e), ptr(cx.cleared), ptr(cx.tet_off),
             ptr(cx.tet_key), ptr(cx.cof2_off), C2, ptr(cx.cof2_data)), (s,), dev)
    TIMER.stop(tok)
    cx.kind2, cx.pivot2, h2, stats = _reduce_and_diagram(
        K=T, col_base=E, off=cx.cof2_off, data=cx.cof2_data, skip=cx.cleared,
        row_cmf=cx.tet_cmf, row_base=E + T, nrows=Q, C=C2, col_edge=cx.tri_e,
        row_edge=cx.tet_e, edge_filt=cx.edge_filt, E=E, ess_ascending=False, device=dev,
        ref_sizes=[(C2, T, Q)] if Q > EXACT_SERIAL_MIN_ROWS else None,
        serial_exact=Q > EXACT_SERIAL_MIN_ROWS, want_edges=cx.want_edges)
    cx.h2 = h2[0]
    if cx.want_edges:
        cx.h2_edges = stats.pop("row_edges")[0]
    cx.info["h2_stats"] = stats


def compute(pts: torch.Tensor, max_edge_length: float, max_dim: int,
            semantics: str = "exact", want_edges: bool = False,
            colex64: bool = False) -> tuple[list, Complex]:
    """Diagrams [H0, ..., H_max_dim] of a CUDA point cloud with n >= 2."""
    cx = build_edges(pts, max_edge_length, colex64=colex64)
    cx.want_edges = want_edges
    build_mst(cx)
    out = [cx.h0]
    if max_dim == 1:
        build_h1(cx, semantics=semantics)
        out.append(cx.h1)
    if max_dim >= 2:
        build_h1(cx, semantics="exact", defer=True)
        build_h2(cx)
        out += [cx.h1, cx.h2]
    return out, cx


# ----------------------------------------------------------------- batches

_KNN_TILE = 256
KNN_MAX_K = 128
_SCREEN_TILE = 64
# Largest cloud whose local colex index fits the 32-bit key field, so that
# keys are unique and the per-cloud order is the key order (rips.py:36-49).
MAX_BATCH_CLOUD = 92682


@dataclass
class BatchLayout:
    """C clouds concatenated on one device (see fph_edge_screen_batch)."""
    C: int
    N: int
    ns: list
    ds: list
    vbase_host: list
    pts: torch.Tensor       # flat float32, cloud c at foff[c]
    foff: torch.Tensor      # int64 (C+1)
    vbase: torch.Tensor     # int64 (C+1)
    tile_off: torch.Tensor  # int64 (C+1)
    qtile_off: torch.Tensor  # int64 (C+1)
    dims: torch.Tensor      # int32 (C)
    n_tiles: int
    n_qtiles: int
    max_d: int


def layout_batch(clouds: list, device) -> BatchLayout:
    """Pack float32 (n_c, d_c) CUDA tensors into one buffer plus offsets."""
    C = len(clouds)
    ns = [int(c.shape[0]) for c in clouds]
    ds = [int(c.shape[1]) for c in clouds]
    ns_a = np.asarray(ns, dtype=np.int64)
    ds_a = np.asarray(ds, dtype=np.int64)
    z = np.zeros(1, dtype=np.int64)
    vbase = np.concatenate([z, np.cumsum(ns_a)])
    foff = np.concatenate([z, np.cumsum(ns_a * ds_a)])
    nt = (ns_a + _SCREEN_TILE - 1) // _SCREEN_TILE
    tile_off = np.concatenate([z, np.cumsum(nt * (nt + 1) // 2)])
    qtile_off = np.concatenate([z, np.cumsum((ns_a + _KNN_TILE - 1) // _KNN_TILE)])
    host = np.concatenate([foff, vbase, tile_off, qtile_off])
    dev_i64 = torch.from_numpy(host).to(device, non_blocking=False)
    parts = [c.reshape(-1) for c in clouds if c.numel() > 0]
    pts = torch.cat(parts) if parts else torch.zeros(1, dtype=torch.float32, device=device)
    k = C + 1
    return BatchLayout(
        C=C, N=int(vbase[-1]), ns=ns, ds=ds, vbase_host=vbase.tolist(), pts=pts,
        foff=dev_i64[0:k], vbase=dev_i64[k:2 * k], tile_off=dev_i64[2 * k:3 * k],
        qtile_off=dev_i64[3 * k:4 * k],
        dims=torch.from_numpy(ds_a.astype(np.int32)).to(device),
        n_tiles=int(tile_off[-1]), n_qtiles=int(qtile_off[-1]), max_d=max(ds) if ds else 1)


def knn_threshold_batch(L: BatchLayout, k: int, percentile: float,
                        want_kth: bool = False):
    """Per-cloud auto_threshold on the device: (t, thr_sq[, kth]) tensors."""
    dev = L.pts.device
    t = _empty(torch.float32, dev, L.C)
    thr_sq = _empty(torch.float32, dev, L.C)
    kth = _empty(torch.float32, dev, max(L.N, 1)) if want_kth else None
    call_ws("fph_knn_threshold_batch",
            (ptr(L.pts), ptr(L.foff), ptr(L.dims), ptr(L.vbase), ptr(L.qtile_off), L.n_qtiles,
             L.C, L.N, int(k), float(percentile), max(1, L.max_d), ptr(kth), ptr(t),
             ptr(thr_sq)), (stream_of(dev),), dev)
    return (t, thr_sq, kth) if want_kth else (t, thr_sq)


def build_edges_batch(L: BatchLayout, thr_sq: torch.Tensor) -> Complex:
    """Stage (a) for a batch: per-cloud screening and filtration order, one
    CSR over all clouds (global vertex ids)."""
    lib = load()
    dev = L.pts.device
    s = stream_of(dev)
    N = L.N
    cx = Complex(n=N, device=dev)
    cx.vbase_host = L.vbase_host
    cap = max(_EDGE_CAP_MIN, _EDGE_CAP_PER_POINT * N)
    cnt = _empty(torch.int64, dev, 1)
    while True:
        keys = _empty(torch.int64, dev, cap)
        vals = _empty(torch.int64, dev, cap)
        tok = TIMER.start("edge_screen", dev)
        check(lib.fph_edge_screen_batch(ptr(L.pts), ptr(L.foff), ptr(L.dims), ptr(L.vbase),
                                        ptr(thr_sq), ptr(L.tile_off), L.n_tiles, L.C,
                                        ptr(keys), ptr(vals), cap, ptr(cnt), s),
              "fph_edge_screen_batch")
        TIMER.stop(tok)
        E = int(cnt.item())
        if E <= cap:
            break
        cap = E
    assert E < 2**31, f"Edge count {E} exceeds int32 range; reduce max_edge_length"
    cx.E = E
    cx.edge_i = _empty(torch.int32, dev, max(E, 1))
    cx.edge_j = _empty(torch.int32, dev, max(E, 1))
    cx.edge_dsq = _empty(torch.float32, dev, max(E, 1))
    cx.edge_filt = _empty(torch.float32, dev, max(E, 1))
    cx.edge_off = _empty(torch.int32, dev, L.C + 1)
    call_ws("fph_edge_sort_batch",
            (ptr(keys), ptr(vals), E, ptr(L.vbase), L.C, ptr(cx.edge_i), ptr(cx.edge_j),
             ptr(cx.edge_dsq), ptr(cx.edge_filt), ptr(cx.edge_off)), (s,), dev)
    del keys, vals
    cx.adj_ptr = _empty(torch.int32, dev, N + 1)
    cx.adj_idx = _empty(torch.int32, dev, max(2 * E, 1))
    cx.adj_rank = _empty(torch.int32, dev, max(2 * E, 1))
    cx.adj_dsq = _empty(torch.float32, dev, max(2 * E, 1))
    call_ws("fph_vertex_csr",
            (ptr(cx.edge_i), ptr(cx.edge_j), ptr(cx.edge_dsq), E, N, ptr(cx.adj_ptr),
             ptr(cx.adj_idx), ptr(cx.adj_rank), ptr(cx.adj_dsq)), (s,), dev)
    return cx


def build_mst_batch(L: BatchLayout, cx: Complex) -> None:
    """Stage (b) for a batch: one forest over the union, H0 per cloud."""
    lib = load()
    dev, N, E = cx.device, cx.n, cx.E
    s = stream_of(dev)
    cx.mst_mask = _empty(torch.uint8, dev, max(E, 1))
    conv = _empty(torch.int32, dev, 1)
    scratch = _lib.WORKSPACE.get(lib.fph_boruvka_scratch_bytes(N), dev)
    check(lib.fph_boruvka_mst(ptr(cx.edge_i), ptr(cx.edge_j), E, N, ptr(cx.mst_mask), None,
                              ptr(conv), ptr(scratch), scratch.numel(), s), "fph_boruvka_mst")
    cx.h0 = _empty(torch.float32, dev, max(N, 1), 2)
    call_ws("fph_h0_diagram_batch",
            (ptr(cx.mst_mask), ptr(cx.edge_dsq), E, ptr(cx.edge_off), ptr(L.vbase), L.C, N,
             ptr(cx.h0)), (s,), dev)
    cx.info["converged"] = conv


def compute_batch(L: BatchLayout, thr_sq: torch.Tensor, max_dim: int,
                  max_triangles: int = MAX_TRIANGLES, semantics: str = "exact",
                  want_edges: bool = False) -> tuple[list, Complex]:
    """Diagrams [[H0, H1?] per cloud] of a packed batch (max_dim <= 1)."""
    if L.N == 0:
        empty = torch.empty(0, 2, dtype=torch.float32, device=L.pts.device)
        return [[empty] * (max_dim + 1) for _ in range(L.C)], None
    cx = build_edges_batch(L, thr_sq)
    cx.want_edges = want_edges
    build_mst_batch(L, cx)
    vb = L.vbase_host
    h0s = [cx.h0[vb[c]:vb[c + 1]] for c in range(L.C)]
    if max_dim == 0:
        if not int(cx.info["converged"].item()):
            raise _lib.FlashPHError("Boruvka did not converge")
        return [[h] for h in h0s], cx
    build_h1(cx, max_triangles, semantics=semantics)
    return [[h0, h1] for h0, h1 in zip(h0s, cx.h1)], cx
