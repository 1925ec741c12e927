"""Seeded point clouds shared by the parity tests.

Inputs are generated with numpy default_rng in FP32 (portable bit for bit);
the families follow the reference's tests (testing/test_reduce_h1.py:59-268,
testing/test_persistence.py:288-306, 476-492, 541-570).
"""
from __future__ import annotations

import numpy as np


def randn(n, d, seed, scale=1.0):
    rng = np.random.default_rng(seed)
    return (rng.standard_normal((n, d)) * scale).astype(np.float32)


def clifford(n, seed, noise=0.0):
    rng = np.random.default_rng(seed)
    th = rng.uniform(0, 2 * np.pi, n).astype(np.float32)
    ph = rng.uniform(0, 2 * np.pi, n).astype(np.float32)
    pts = np.column_stack([np.cos(th), np.sin(th), np.cos(ph), np.sin(ph)]).astype(
        np.float32) / np.sqrt(2)
    if noise:
        pts = pts + rng.normal(0, noise, pts.shape)
    return pts.astype(np.float32)


def o3(n, seed):
    rng = np.random.default_rng(seed)
    mats = rng.standard_normal((n, 3, 3))
    out = np.empty((n, 9), dtype=np.float32)
    for i in range(n):
        q, r = np.linalg.qr(mats[i])
        dd = np.sign(np.diag(r))
        dd[dd == 0] = 1.0
        out[i] = (q * dd[np.newaxis, :]).ravel()
    return out


def sphere(n, d, seed, noise=0.0):
    rng = np.random.default_rng(seed)
    x = rng.standard_normal((n, d))
    x /= np.linalg.norm(x, axis=1, keepdims=True)
    if noise:
        x = x + rng.normal(0, noise, x.shape)
    return x.astype(np.float32)


def circle(n):
    th = np.linspace(0, 2 * np.pi, n, endpoint=False).astype(np.float32)
    return np.column_stack([np.cos(th), np.sin(th)]).astype(np.float32)


def grid3():
    return np.array([[x, y, z] for x in range(3) for y in range(3) for z in range(3)],
                    dtype=np.float32)


# (name, points, max_edge_length, max_dim)
SMALL = [
    ("randn_d2_s42", lambda: randn(30, 2, 42), 1.5, 2),
    ("randn_d2_s43", lambda: randn(30, 2, 43), 1.5, 2),
    ("randn_d2_s44", lambda: randn(30, 2, 44), 1.5, 2),
    ("randn_d4_s42", lambda: randn(25, 4, 42), 2.0, 2),
    ("randn_d10_s42", lambda: randn(20, 10, 42), 3.0, 2),
    ("randn_d3_n100", lambda: randn(100, 3, 42), 1.5, 2),
    ("clifford_200", lambda: clifford(200, 42), 0.6, 2),
    ("o3_200", lambda: o3(200, 42), 1.8, 1),
    ("grid3_ties", grid3, 2.0, 2),
    ("circle_30", lambda: circle(30), 1.5, 2),
    ("circle_20_small_t", lambda: circle(20), 0.5, 1),
    ("cluster_tight", lambda: randn(10, 2, 42, 0.01), 0.5, 2),
    ("no_edges", lambda: randn(50, 3, 7), 1e-10, 2),
    ("sphere2_noisy", lambda: sphere(60, 3, 5, 0.05), 1.2, 2),
    ("d50_n30", lambda: randn(30, 50, 3), 10.0, 1),
]

MEDIUM = [
    ("sphere3_192", lambda: sphere(192, 4, 42), 2.001, 1),
    ("o3_1024", lambda: o3(1024, 42), 1.8, 1),
    ("randn_d10_n2000", lambda: randn(2000, 10, 1), 2.2, 2),
    ("clifford_1000_noisy", lambda: clifford(1000, 3, 0.02), 0.45, 2),
    ("randn_d10_n10000_t16", lambda: randn(10000, 10, 0), 1.6, 2),
]


def batch_uneven(C, seed, nmin=200, nmax=5000, dims=(4, 10, 50), noise=0.05):
    """Clouds of log-uniform size in [nmin, nmax] and dimension drawn from
    ``dims``; points uniform on the unit (d-1)-sphere plus N(0, noise^2)."""
    rng = np.random.default_rng(seed)
    ns = np.exp(rng.uniform(np.log(nmin), np.log(nmax), C)).astype(np.int64)
    ns = np.clip(ns, nmin, nmax)
    ds = rng.choice(np.asarray(dims), C)
    clouds = []
    for n, d in zip(ns, ds):
        x = rng.standard_normal((int(n), int(d)))
        x /= np.linalg.norm(x, axis=1, keepdims=True)
        x = x + rng.normal(0, noise, x.shape)
        clouds.append(x.astype(np.float32))
    return clouds


def clifford_batch(C, n, seed, noise=0.02):
    """C Clifford-torus clouds (cos t, sin t, cos p, sin p)/sqrt(2) + noise."""
    return [clifford(n, seed + c, noise) for c in range(C)]


def grid_points(m: int) -> np.ndarray:
    """m^3 integer grid points in 3-D, x-major (exact in FP32, so every
    squared distance is an exact small integer: massive dsq ties)."""
    return np.ascontiguousarray(np.indices((m, m, m)).reshape(3, -1).T, dtype=np.float32)


def grid_edges(pts: np.ndarray, t: float):
    """(i, j, dsq) of every pair with dsq < fp32(t^2) on integer points, from
    scipy's k-d tree (FP64 distances are exact here)."""
    from scipy.spatial import cKDTree
    pr = cKDTree(pts.astype(np.float64)).query_pairs(float(t) + 0.01, output_type="ndarray")
    i = pr.min(axis=1).astype(np.int64)
    j = pr.max(axis=1).astype(np.int64)
    dsq = ((pts[i].astype(np.float64) - pts[j]) ** 2).sum(axis=1).astype(np.float32)
    keep = dsq < np.float32(float(t) * float(t))
    return i[keep].astype(np.int32), j[keep].astype(np.int32), dsq[keep]


def reference_edge_order(i, j, dsq) -> np.ndarray:
    """Permutation putting edges into the reference's filtration order,
    written with numpy: emission order first -- tile (i // 32, j // 32)
    row-major, then (i mod 32, j mod 32) (kernels/edge_kernel.py:133-149,
    199-210) -- then a stable sort on the packed key
    float_bits(dsq) << 32 | (j (j - 1) / 2 + i) mod 2^32 (rips.py:41-49,
    189-195)."""
    i = np.asarray(i, dtype=np.int64)
    j = np.asarray(j, dtype=np.int64)
    em = np.lexsort((j % 32, i % 32, j // 32, i // 32))
    bits = np.asarray(dsq, dtype=np.float32).view(np.uint32).astype(np.int64)
    key = (bits << 32) | ((j * (j - 1) // 2 + i) & 0xFFFFFFFF)
    return em[np.argsort(key[em], kind="stable")]


def wrapped_key_ties(i, j, dsq) -> int:
    """Number of edges whose packed key equals another edge's (possible
    only once the colex index passes 2^32, i.e. n > 92,682)."""
    i = np.asarray(i, dtype=np.int64)
    j = np.asarray(j, dtype=np.int64)
    bits = np.asarray(dsq, dtype=np.float32).view(np.uint32).astype(np.int64)
    key = (bits << 32) | ((j * (j - 1) // 2 + i) & 0xFFFFFFFF)
    _, c = np.unique(key, return_counts=True)
    return int(c[c > 1].sum())


def near_threshold_pairs(m: int, d: int, t: float, seed: int, spread: int = 4):
    """2m points: m N(0,1) points x_p and partners y_p placed so that the FP32
    FMA-chain dsq(x_p, y_p) lies within a few ulp of fp32(t^2) on either side
    (target offsets uniform in [-spread, spread] ulp, reached by rescaling
    y_p - x_p a few times).  Returns (points, achieved ulp offsets)."""
    import oracle
    rng = np.random.default_rng(seed)
    x = rng.standard_normal((m, d)).astype(np.float32)
    v = rng.standard_normal((m, d))
    v /= np.linalg.norm(v, axis=1, keepdims=True)
    t2 = np.float32(float(t) * float(t))
    t2_bits = int(t2.view(np.int32))
    targets = t2_bits + rng.integers(-spread, spread + 1, m)
    y = np.empty_like(x)
    offs = np.empty(m, dtype=np.int64)
    for p in range(m):
        s = float(t)
        tgt = float(np.array(targets[p], dtype=np.int32).view(np.float32))
        for _ in range(8):
            yp = (x[p].astype(np.float64) + s * v[p]).astype(np.float32)
            dsq = oracle.dist_sq(np.stack([x[p], yp]), 0, 1)
            if float(dsq) == tgt:
                break
            s *= np.sqrt(tgt / float(dsq))
        y[p] = yp
        offs[p] = int(np.float32(dsq).view(np.int32)) - t2_bits
    return np.concatenate([x, y]).astype(np.float32), offs
