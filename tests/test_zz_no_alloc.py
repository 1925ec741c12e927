"""fph_boruvka_mst_noalloc (the allocation-free Boruvka entry point) against
fph_boruvka_mst, called through the C ABI: identical forest mask, component
labels and convergence flag on two calls in a row, the second on a smaller
complex, both large enough for the multi-launch form that uses the scratch
array, plus one call small enough for the one-workgroup form."""
from __future__ import annotations

import ctypes

import numpy as np
import pytest
import torch

from flash_ph import _lib
from flash_ph.rips import rips_filtration

pytestmark = pytest.mark.gpu


def _edges(n, d, t, seed):
    rng = np.random.default_rng(seed)
    pts = torch.from_numpy(rng.standard_normal((n, d)).astype(np.float32)).cuda()
    f = rips_filtration(pts, t)
    return f.edge_i.contiguous(), f.edge_j.contiguous()


def _forest(lib, ei, ej, n, noalloc):
    E = ei.numel()
    dev = ei.device
    mask = torch.full((max(E, 1),), 7, dtype=torch.uint8, device=dev)
    lab = torch.full((n,), -5, dtype=torch.int64, device=dev)
    conv = torch.full((1,), -1, dtype=torch.int32, device=dev)
    s = _lib.stream_of(dev)
    if noalloc:
        rc = lib.fph_boruvka_mst_noalloc(ei.data_ptr(), ej.data_ptr(), E, n, mask.data_ptr(),
                                         lab.data_ptr(), conv.data_ptr(), s)
    else:
        nb = lib.fph_boruvka_scratch_bytes(n)
        scratch = torch.empty(nb, dtype=torch.uint8, device=dev)
        rc = lib.fph_boruvka_mst(ei.data_ptr(), ej.data_ptr(), E, n, mask.data_ptr(),
                                 lab.data_ptr(), conv.data_ptr(), scratch.data_ptr(), nb, s)
    assert rc == 0, rc
    torch.cuda.synchronize()
    return mask[:E].cpu().numpy(), lab.cpu().numpy(), int(conv.item())


# (n, d, t, seed): 40,000 and then 15,000 points take the multi-launch form
# (n > 12,288); 3,000 points the one-workgroup form.
CALLS = [(40_000, 3, 0.12, 1), (15_000, 3, 0.16, 2), (3_000, 2, 0.05, 3)]


def test_noalloc_matches_scratch_entry_point():
    lib = _lib.load()
    for n, d, t, seed in CALLS:
        print(f"noalloc boruvka n={n} d={d} t={t} seed={seed}", flush=True)
        ei, ej = _edges(n, d, t, seed)
        a = _forest(lib, ei, ej, n, noalloc=True)
        b = _forest(lib, ei, ej, n, noalloc=False)
        assert a[2] == 1 and b[2] == 1
        assert np.array_equal(a[0], b[0])
        assert np.array_equal(a[1], b[1])
        # a forest: n - #components edges
        assert int(a[0].sum()) == n - len(np.unique(a[1]))


def test_noalloc_rejects_oversized_input():
    lib = _lib.load()
    n = 262_145
    rc = lib.fph_boruvka_mst_noalloc(None, None, 0, n, None, None, None, None)
    assert rc == _lib.FPH_ERR_LIMIT
