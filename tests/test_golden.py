"""Golden fixtures (tests/golden/*.npz): the oracle reproduces them on the
CPU, and the HIP pipeline reproduces them bit for bit on the GPU."""
from __future__ import annotations

import glob
import os

import numpy as np
import pytest
import torch

import oracle

FILES = sorted(glob.glob(os.path.join(os.path.dirname(__file__), "golden", "*.npz")))
IDS = [os.path.basename(f)[:-4] for f in FILES]


def _load(path):
    with np.load(path, allow_pickle=False) as z:
        return {k: z[k] for k in z.files}


def _same(a, b):
    a = np.ascontiguousarray(a)
    b = np.ascontiguousarray(b)
    if a.dtype == np.float32:
        a, b = a.view(np.uint32), b.view(np.uint32)
    return a.shape == b.shape and np.array_equal(a, b)


@pytest.mark.parametrize("path", FILES, ids=IDS)
def test_oracle_reproduces_fixture(path):
    g = _load(path)
    md = int(g["max_dim"])
    r = oracle.run(g["points"], float(g["max_edge_length"]), md)
    for key in ("edge_i", "edge_j", "edge_dsq", "mst_idx"):
        assert _same(r[key], g[key]), key
    for k in range(md + 1):
        assert _same(r["diagrams"][k], g[f"h{k}"]), f"h{k}"


@pytest.mark.gpu
@pytest.mark.parametrize("path", FILES, ids=IDS)
def test_gpu_reproduces_fixture(path):
    from flash_ph import rips_persistence
    g = _load(path)
    md = int(g["max_dim"])
    print(f"[gpu] golden {os.path.basename(path)}", flush=True)
    dg = rips_persistence(torch.from_numpy(g["points"]).cuda(), float(g["max_edge_length"]),
                          max_dim=md)
    for k in range(md + 1):
        assert _same(dg[k].cpu().numpy(), g[f"h{k}"]), f"h{k}"
