"""Autotuner for libflashph's launch parameters.

Four knobs set how work is spread over the GPU and never what is computed:

* ``screen_tile``: points per side of a distance tile of the direct edge
  screen (32, 64, 128), per (n, d) bucket;
* ``enum_block``: threads per workgroup of the link kernels (triangles,
  tetrahedra, cofacet columns; one wavefront per edge or triangle), per
  item-count bucket;
* ``apparent_block``: columns per workgroup of the apparent-pair and
  clearing kernels, per column-count bucket;
* ``queue_wgs``: workgroups of the residual reduction's queue (1: one
  workgroup; more: the device-wide queue), per column-count bucket.

The library keeps one value per bucket (floor(log2 size), floor(log2 d)).
``tune`` times every candidate on the caller's own cloud with HIP events on
the launch stream (median of ``repeats``), checks that every candidate
returns the same diagrams, installs the fastest and, with ``persist``, adds
it to a JSON cache that ``load_cache`` reapplies in a later process (the file
named by ``FLASH_PH_TUNE_CACHE``, else ``~/.cache/flash_ph/autotune-gfx950.json``).
The reference hard-codes its Triton block sizes (kernels/edge_kernel.py:193-222,
kernels/triangle_kernel.py:366-397); this module replaces those constants.
"""
from __future__ import annotations

import ctypes
import json
import os
import statistics

import torch

from flash_ph import _lib

KNOBS = {"screen_tile": 0, "enum_block": 1, "apparent_block": 2, "queue_wgs": 3}


def _knob_id(knob) -> int:
    if isinstance(knob, int) and knob in KNOBS.values():
        return knob
    if knob in KNOBS:
        return KNOBS[knob]
    raise ValueError(f"unknown autotune knob {knob!r}; expected one of {sorted(KNOBS)}")


def candidates(knob) -> list:
    lib = _lib.load()
    buf = (ctypes.c_int32 * 16)()
    m = lib.fph_autotune_candidates(_knob_id(knob), ctypes.addressof(buf), 16)
    _lib.check(min(m, 0), "fph_autotune_candidates")
    return [int(buf[i]) for i in range(m)]


def query(knob, size: int, d: int = 0) -> int:
    v = _lib.load().fph_autotune_query(_knob_id(knob), int(size), int(d))
    _lib.check(min(v, 0), "fph_autotune_query")
    return int(v)


def set_value(knob, size: int, d: int, value: int) -> None:
    """Installs `value` (0: the default) for the bucket of (size, d)."""
    _lib.check(_lib.load().fph_autotune_set(_knob_id(knob), int(size), int(d), int(value)),
               "fph_autotune_set")


def reset() -> None:
    _lib.load().fph_autotune_reset()


def cache_path() -> str:
    p = os.environ.get("FLASH_PH_TUNE_CACHE")
    if p:
        return p
    return os.path.join(os.path.expanduser("~"), ".cache", "flash_ph", "autotune-gfx950.json")


def _read_cache(path: str) -> list:
    if not os.path.exists(path):
        return []
    with open(path) as f:
        data = json.load(f)
    return list(data.get("entries", []))


def load_cache(path: str | None = None) -> int:
    """Applies every cached entry; returns how many were applied."""
    k = 0
    for e in _read_cache(path or cache_path()):
        set_value(e["knob"], e["size"], e["d"], e["value"])
        k += 1
    return k


def save_cache(entries: list, path: str | None = None) -> str:
    """Merges `entries` (dicts with knob, size, d, value) into the cache,
    replacing entries of the same knob and bucket."""
    path = path or cache_path()

    def bucket(e):
        b = lambda x: max(int(x), 1).bit_length() - 1
        return (e["knob"], b(e["size"]), b(e["d"]))

    merged = {bucket(e): e for e in _read_cache(path)}
    for e in entries:
        merged[bucket(e)] = e
    os.makedirs(os.path.dirname(os.path.abspath(path)), exist_ok=True)
    tmp = path + ".tmp"
    with open(tmp, "w") as f:
        json.dump({"arch": "gfx950", "entries": list(merged.values())}, f, indent=1)
    os.replace(tmp, path)
    return path


def _run(pts, max_edge_length, max_dim, semantics):
    from flash_ph import _pipeline
    dgms, cx = _pipeline.compute(pts, max_edge_length, max_dim, semantics=semantics)
    return dgms, cx


def _timed(stage, pts, max_edge_length, max_dim, semantics, repeats):
    """Median over `repeats` pipeline runs of the summed event time of
    `stage`, plus the diagrams of the last run."""
    from flash_ph import _pipeline
    times, dgms = [], None
    for _ in range(repeats):
        _pipeline.TIMER.enable(True)
        try:
            dgms, _ = _run(pts, max_edge_length, max_dim, semantics)
            summ = _pipeline.TIMER.summary()
        finally:
            _pipeline.TIMER.enable(False)
        cnt, mean = summ.get(stage, (0, 0.0))
        times.append(cnt * mean)
    return statistics.median(times), dgms


def _same(a, b) -> bool:
    return len(a) == len(b) and all(torch.equal(x, y) for x, y in zip(a, b))


def tune(points, max_edge_length: float, max_dim: int = 1, repeats: int = 5,
         knobs=None, semantics: str = "reference", persist: bool = False,
         path: str | None = None) -> dict:
    """Times every candidate of each knob on this cloud and installs the
    fastest.  Returns {knob: {"value", "buckets", "ms": {candidate: ms}}}.
    Raises RuntimeError if two candidates give different diagrams."""
    from flash_ph import _pipeline
    _lib.require_gpu()
    pts = torch.as_tensor(points)
    if pts.device.type != "cuda":
        pts = pts.cuda()
    pts = pts.float().contiguous()
    n, d = pts.shape
    knobs = list(knobs or KNOBS)
    ref, cx = _run(pts, max_edge_length, max_dim, semantics)       # warm-up and sizes
    E, T, K2 = cx.E, cx.T, cx.T
    keys = {
        "screen_tile": [(n, d)],
        "enum_block": [(E, 0)] + ([(T, 0)] if max_dim >= 2 else []),
        "apparent_block": [(E, 0)] + ([(K2, 0)] if max_dim >= 2 else []),
        "queue_wgs": [(E, 0)] + ([(K2, 0)] if max_dim >= 2 else []),
    }
    stage = {"screen_tile": "edge_screen", "enum_block": "enumerate",
             "apparent_block": "apparent", "queue_wgs": "reduce"}
    out, entries = {}, []
    for knob in knobs:
        _knob_id(knob)
        if knob == "screen_tile" and _pipeline.use_mfma_screen(n, d):
            continue          # the matrix-core screen has no tile knob
        if knob != "screen_tile" and max_dim < 1:
            continue
        before = [query(knob, s_, d_) for s_, d_ in keys[knob]]
        ms = {}
        try:
            for v in candidates(knob):
                for s_, d_ in keys[knob]:
                    set_value(knob, s_, d_, v)
                t, dg = _timed(stage[knob], pts, max_edge_length, max_dim, semantics, repeats)
                if not _same(dg, ref):
                    raise RuntimeError(f"autotune: {knob}={v} changed the diagrams")
                ms[v] = round(t, 5)
        except BaseException:
            for (s_, d_), v in zip(keys[knob], before):
                set_value(knob, s_, d_, v)
            raise
        best = min(ms, key=ms.get)
        for s_, d_ in keys[knob]:
            set_value(knob, s_, d_, best)
            entries.append({"knob": knob, "size": int(s_), "d": int(d_), "value": int(best)})
        out[knob] = {"value": best, "buckets": keys[knob], "ms": ms}
    if persist and entries:
        save_cache(entries, path)
    return out
