"""Cofacet columns sorted by either method -- the one-launch comparison sort
(columns of up to 256 values in registers, longer ones against memory) or the
segmented radix sort -- equal the oracle's columns bit for bit.  The library
picks the method from the columns' sizes; FLASH_PH_COLSORT_WORK forces one."""
import numpy as np
import pytest
import torch

import cases
import oracle
from flash_ph import _pipeline

pytestmark = pytest.mark.gpu

# forced: "0" = always the radix sort, "1e30" = always the comparison sort
SORTS = ["0", "1e30"]


def _check_h1_columns(cx, ref):
    E = cx.E
    off = cx.cof_off.cpu().numpy()
    dat = cx.cof_data.cpu().numpy().view(np.uint32)
    roff, rdat = ref["cof_off"], ref["cof_data"].view(np.uint32)
    forest = cx.mst_mask.cpu().numpy()[:E].astype(bool)
    lo, lr = off[1:] - off[:-1], roff[1:] - roff[:-1]
    assert (lo[forest] == 0).all()
    assert np.array_equal(lo[~forest], lr[~forest])
    keep = np.repeat(~forest, lr)           # the oracle's entries of non-forest columns
    assert np.array_equal(dat[:off[-1]], rdat[keep])
    return int(lr.max())


@pytest.mark.parametrize("work", SORTS)
def test_long_columns_match_oracle(monkeypatch, work):
    # n=600 in 2-D at t=1.2: columns up to 342 values (past the 256 held in
    # registers), C^2 / K = 1.6e9 (the default picks the radix sort here)
    print(f"[gpu] colsort long work={work}", flush=True)
    monkeypatch.setenv("FLASH_PH_COLSORT_WORK", work)
    pts = cases.randn(600, 2, 7)
    ref = oracle.run(pts, 1.2, 1)
    dg, cx = _pipeline.compute(torch.from_numpy(pts).cuda(), 1.2, 1, semantics="reference")
    assert _check_h1_columns(cx, ref) > 256
    assert np.array_equal(dg[1].cpu().numpy().view(np.uint32), ref["h1"].view(np.uint32))


@pytest.mark.parametrize("work", SORTS)
def test_h1_h2_columns_match_oracle(monkeypatch, work):
    print(f"[gpu] colsort h2 work={work}", flush=True)
    monkeypatch.setenv("FLASH_PH_COLSORT_WORK", work)
    pts = cases.randn(400, 3, 11)
    ref = oracle.run(pts, 1.0, 2)
    dg, cx = _pipeline.compute(torch.from_numpy(pts).cuda(), 1.0, 2)
    _check_h1_columns(cx, ref)
    for k in (0, 1, 2):   # rows in order, bit for bit
        a, b = dg[k].cpu().numpy(), ref["diagrams"][k]
        assert a.shape == b.shape and np.array_equal(a.view(np.uint32), b.view(np.uint32)), f"H{k}"
