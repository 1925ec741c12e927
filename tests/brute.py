"""Independent checkers for the oracle (CPU, small inputs only).

Nothing here shares code with oracle/fph_oracle.c:
* ``fma_chain_exact``: the FP32 distance chain with every fused multiply-add
  evaluated in exact rational arithmetic and rounded once to FP32
  (round-to-nearest-even), i.e. the IEEE-754 definition of fmaf.
* ``brute_diagrams``: the full Vietoris-Rips complex up to dimension
  max_dim + 1 by clique enumeration, reduced with the textbook *homology*
  column algorithm (boundary matrix, pivot = lowest row, no clearing, no
  apparent pairs, no cohomology) over Z/2 with Python integers as bitsets.
  Persistence diagrams do not depend on the reduction strategy, so this pins
  the oracle's cohomology-with-shortcuts diagrams as multisets.
"""
from __future__ import annotations

import itertools
from fractions import Fraction

import numpy as np


def _f32_bits_even(x: np.float32) -> bool:
    return (int(np.array(x, dtype=np.float32).view(np.uint32)) & 1) == 0


def round_f32(fr: Fraction) -> np.float32:
    """Exact rational -> nearest FP32, ties to even."""
    if fr == 0:
        return np.float32(0.0)
    f = np.float32(float(fr))
    cands = [np.nextafter(f, np.float32(-np.inf)), f, np.nextafter(f, np.float32(np.inf))]
    best = None
    for c in cands:
        if not np.isfinite(c):
            continue
        dist = abs(Fraction(float(c)) - fr)
        if best is None or dist < best[0] or (dist == best[0] and _f32_bits_even(c)):
            best = (dist, c)
    return np.float32(best[1])


def fmaf_exact(a, b, c) -> np.float32:
    return round_f32(Fraction(float(a)) * Fraction(float(b)) + Fraction(float(c)))


def fma_chain_exact(xi, xj) -> np.float32:
    acc = np.float32(0.0)
    for a, b in zip(np.asarray(xi, np.float32), np.asarray(xj, np.float32)):
        diff = np.float32(a - b)           # IEEE FP32 subtraction
        acc = fmaf_exact(diff, diff, acc)
    return acc


def all_dsq(pts) -> np.ndarray:
    pts = np.asarray(pts, np.float32)
    n = pts.shape[0]
    D = np.zeros((n, n), np.float32)
    for i in range(n):
        for j in range(i + 1, n):
            D[i, j] = D[j, i] = fma_chain_exact(pts[i], pts[j])
    return D


def rips_cliques(D, thr_sq, max_k):
    """All simplices (sorted vertex tuples) of dimension <= max_k with every
    edge strictly below thr_sq, with their filtration value (max edge dsq)."""
    n = D.shape[0]
    adj = [[j for j in range(n) if j != i and D[i, j] < thr_sq] for i in range(n)]
    simplices = [((v,), np.float32(0.0)) for v in range(n)]
    frontier = [(v,) for v in range(n)]
    for k in range(1, max_k + 1):
        nxt = []
        for s in frontier:
            for w in adj[s[-1]]:
                if w <= s[-1]:
                    continue
                if all(D[u, w] < thr_sq for u in s):
                    t = s + (w,)
                    val = max(D[a, b] for a, b in itertools.combinations(t, 2))
                    simplices.append((t, np.float32(val)))
                    nxt.append(t)
        frontier = nxt
    return simplices


def brute_diagrams(pts, thr_sq, max_dim):
    """Diagrams [H0..H_max_dim] (values = sqrt of FP32 dsq) by homology
    reduction.  H0 keeps every finite pair (zero length included, as the
    reference's H0); higher dimensions keep pairs with death > birth."""
    D = all_dsq(pts)
    simp = rips_cliques(D, np.float32(thr_sq), max_dim + 1)
    simp.sort(key=lambda sv: (float(sv[1]), len(sv[0]), sv[0]))
    index = {s: k for k, (s, _) in enumerate(simp)}
    low_owner = {}
    cols = {}
    paired_birth = set()
    pairs = []
    for j, (s, val) in enumerate(simp):
        col = 0
        if len(s) > 1:
            for f in itertools.combinations(s, len(s) - 1):
                col ^= 1 << index[f]
        while col:
            low = col.bit_length() - 1
            o = low_owner.get(low)
            if o is None:
                break
            col ^= cols[o]
        if col:
            low = col.bit_length() - 1
            low_owner[low] = j
            cols[j] = col
            paired_birth.add(low)
            pairs.append((low, j))
    out = [[] for _ in range(max_dim + 1)]
    for b, d in pairs:
        dim = len(simp[b][0]) - 1
        if dim > max_dim:
            continue
        vb, vd = simp[b][1], simp[d][1]
        if dim == 0 or vd > vb:
            out[dim].append((np.sqrt(vb), np.sqrt(vd)))
    killers = {d for _, d in pairs}
    for k, (s, val) in enumerate(simp):
        dim = len(s) - 1
        if dim <= max_dim and k not in paired_birth and k not in killers:
            out[dim].append((np.sqrt(val), np.float32(np.inf)))
    return [np.array(sorted(o, key=lambda r: (float(r[0]), float(r[1]))), dtype=np.float32)
            .reshape(-1, 2) for o in out]


def multiset(dgm) -> np.ndarray:
    """Rows sorted by (birth, death): order-free comparison of a diagram."""
    a = np.asarray(dgm, np.float32).reshape(-1, 2)
    if a.shape[0] == 0:
        return a
    idx = np.lexsort((a[:, 1], a[:, 0]))
    return a[idx]


def reference_reduction(cof_off, cof_data, skip, r2f, max_col, max_pool):
    """The reference's capacity-limited H1 reduction, written from its
    source with Python containers (a dict for its hash map, lists for its
    pool): apparent pre-detection and pool prefill (reduce_h1_gpu.py:159-223,
    282-289; _numba_utils.py:68-104), then residual columns in descending
    rank order, merged with the MAX_COL-limited sorted XOR
    (_numba_utils.py:43-65) and stored only while the pool has room
    (cohomology_general.py:255-366).  Columns are ranks 0..K-1 (edges),
    rows are global ranks.  Returns (pairs [(birth, death)] in emission
    order, essentials, drops, truncated merges)."""
    K = len(cof_off) - 1
    cols = [list(cof_data[cof_off[k]:cof_off[k + 1]]) for k in range(K)]
    cmf = {}
    for k in range(K):
        if not skip[k]:
            for r in cols[k]:
                cmf[r] = max(cmf.get(r, -1), k)
    total = cof_off[K]
    cap = min(max(2 * total, 16 * K, 100_000), max_pool)
    store = {}        # pivot row -> stored column
    used = 0
    pairs, ess = [], []
    drops = truncs = 0
    apparent = [not skip[k] and len(cols[k]) > 0 and cmf.get(cols[k][0]) == k
                for k in range(K)]
    for k in range(K):
        if apparent[k]:
            if r2f[cols[k][0]] > r2f[k]:
                pairs.append((k, cols[k][0]))
    for k in range(K):
        if apparent[k]:
            if used + len(cols[k]) <= cap and len(store) < K:
                store[cols[k][0]] = list(cols[k])
                used += len(cols[k])

    def xor(a, b):
        out, ia, ib = [], 0, 0
        while ia < len(a) and ib < len(b):
            if len(out) >= max_col:
                break
            if a[ia] < b[ib]:
                out.append(a[ia]); ia += 1
            elif a[ia] > b[ib]:
                out.append(b[ib]); ib += 1
            else:
                ia += 1; ib += 1
        # the tails, a's remainder first, each while room remains
        tail = a[ia:] + b[ib:]
        room = max_col - len(out)
        out += tail[:max(room, 0)]
        return out, len(tail) > max(room, 0)

    for k in range(K - 1, -1, -1):
        if skip[k] or apparent[k]:
            continue
        if not cols[k]:
            ess.append(k)
            continue
        if len(cols[k]) > max_col:
            truncs += 1
        cur = cols[k][:max_col]
        while cur and cur[0] in store:
            cur, cut = xor(cur, store[cur[0]])
            truncs += cut
        if cur:
            if used + len(cur) <= cap and len(store) < K:
                store[cur[0]] = list(cur)
                used += len(cur)
            else:
                drops += 1
            if r2f[cur[0]] > r2f[k]:
                pairs.append((k, cur[0]))
        else:
            ess.append(k)
    return pairs, ess, drops, truncs
