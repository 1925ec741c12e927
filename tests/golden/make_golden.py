"""Writes tests/golden/*.npz: inputs and expected outputs of the known-answer
workloads, computed by the CPU oracle (oracle/) and cross-checked here by the
brute-force homology reduction of tests/brute.py wherever it is small enough.

The reference ships no stored vectors (its tests call ripser at run time), so
these fixtures hold the oracle's answers on the reference's own test inputs
regenerated with numpy default_rng (portable bit for bit).

    python tests/golden/make_golden.py
"""
from __future__ import annotations

import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path[:0] = [os.path.join(ROOT, "oracle"), os.path.join(ROOT, "tests")]

import brute  # noqa: E402
import cases  # noqa: E402
import oracle  # noqa: E402

FIXTURES = [
    # name, points, t, max_dim, brute-check
    ("circle_20_t05", cases.circle(20), 0.5, 1, True),
    ("randn_d2_n30_s42", cases.randn(30, 2, 42), 1.5, 2, False),
    ("randn_d4_n16_s3", cases.randn(16, 4, 3), 2.0, 2, True),
    ("grid3_t2", cases.grid3(), 2.0, 2, False),
    ("clifford_200_t14", cases.clifford(200, 42), 1.4, 2, False),
    ("sphere_40_t2", cases.sphere(40, 3, 42, 0.05), 2.0, 2, False),
    ("o3_200_t18", cases.o3(200, 42), 1.8, 1, False),
]


def main():
    for name, pts, t, md, check in FIXTURES:
        r = oracle.run(pts, t, md)
        if check:
            b = brute.brute_diagrams(pts, oracle.threshold_sq_f32(t), md)
            for k in range(md + 1):
                a, c = brute.multiset(r["diagrams"][k]), brute.multiset(b[k])
                assert a.shape == c.shape and np.array_equal(a.view(np.uint32), c.view(np.uint32))
        out = {"points": pts, "max_edge_length": np.float64(t), "max_dim": np.int64(md),
               "edge_i": r["edge_i"], "edge_j": r["edge_j"], "edge_dsq": r["edge_dsq"],
               "mst_idx": r["mst_idx"]}
        for k in range(md + 1):
            out[f"h{k}"] = r["diagrams"][k]
        np.savez_compressed(os.path.join(HERE, f"{name}.npz"), **out)
        print(name, [d.shape for d in r["diagrams"]])


if __name__ == "__main__":
    main()
