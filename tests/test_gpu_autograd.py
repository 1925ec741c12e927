"""Gradients of diagrams with respect to the points: the HIP backward
against torch autograd on the recomputed critical edge lengths (float64),
against central finite differences (h = 1e-3, float64) of the same
fixed-pairing loss, and against finite differences through the whole
pipeline where the pairing is locally constant."""
from __future__ import annotations

import numpy as np
import pytest
import torch

import cases

pytestmark = pytest.mark.gpu


def _critical_loss64(x64, dgms_edges, weights, edge_i, edge_j):
    """sum over diagram entries of weight * |x_u - x_v| for the entry's edge,
    in float64 (the fixed-pairing form of the loss)."""
    total = torch.zeros((), dtype=torch.float64, device=x64.device)
    for re, w in zip(dgms_edges, weights):
        re = re.reshape(-1).long()
        w = w.reshape(-1).to(torch.float64)
        keep = re >= 0
        e = re[keep]
        u, v = edge_i[e].long(), edge_j[e].long()
        length = torch.linalg.norm(x64[u] - x64[v], dim=1)
        total = total + (w[keep] * length).sum()
    return total


def _forward_with_edges(x, t, max_dim):
    from flash_ph.autograd import RipsPersistenceFunction
    out = RipsPersistenceFunction.apply(x, t, max_dim, "reference")
    return out


@pytest.mark.parametrize("max_dim,seed", [(1, 0), (1, 1), (2, 2)])
def test_backward_matches_torch_autograd_and_fd(max_dim, seed):
    from flash_ph import rips_persistence
    print(f"[gpu] autograd max_dim={max_dim} seed={seed}", flush=True)
    x0 = cases.clifford(300, 40 + seed, noise=0.02)
    t = 0.7
    x = torch.from_numpy(x0).cuda().requires_grad_(True)
    dgms = rips_persistence(x, t, max_dim=max_dim)
    plain = rips_persistence(torch.from_numpy(x0).cuda(), t, max_dim=max_dim)
    for a, b in zip(dgms, plain):   # same values with and without grad
        assert torch.equal(a.detach().view(torch.int32), b.view(torch.int32))
    # random upstream weights on every finite entry
    g = torch.Generator(device="cpu").manual_seed(seed)
    weights = [torch.randn(d.shape, generator=g).cuda() for d in dgms]
    weights = [torch.where(torch.isfinite(d.detach()), w, torch.zeros_like(w))
               for d, w in zip(dgms, weights)]
    loss = sum((torch.where(torch.isfinite(d), d, torch.zeros_like(d)) * w).sum()
               for d, w in zip(dgms, weights))
    loss.backward()
    ours = x.grad.double()

    # the critical edges recorded by the forward pass
    from flash_ph import autograd as fag
    from flash_ph import _pipeline
    fn_out, cx = _pipeline.compute(torch.from_numpy(x0).cuda(), t, max_dim,
                                   semantics="exact" if max_dim == 2 else "reference",
                                   want_edges=True)
    eo = torch.tensor([0, cx.E], dtype=torch.int32, device="cuda")
    edges = [fag._h0_edges(cx.mst_mask, cx.E, eo, [0, 300], 300, x.device)]
    edges.append(cx.h1_edges)
    if max_dim == 2:
        edges.append(cx.h2_edges)
    x64 = torch.from_numpy(x0).double().cuda().requires_grad_(True)
    ref_loss = _critical_loss64(x64, edges, weights, cx.edge_i, cx.edge_j)
    ref_loss.backward()
    ref = x64.grad
    assert abs(ref_loss.item() - loss.item()) <= 1e-4 * max(1.0, abs(ref_loss.item()))
    assert torch.allclose(ours, ref, rtol=1e-4, atol=1e-5), (ours - ref).abs().max().item()

    # central differences of the fixed-pairing loss, float64, h = 1e-3
    h = 1e-3
    rng = np.random.default_rng(seed)
    for _ in range(12):
        i, k = int(rng.integers(300)), int(rng.integers(4))
        xp = torch.from_numpy(x0).double().cuda()
        xm = xp.clone()
        xp[i, k] += h
        xm[i, k] -= h
        with torch.no_grad():
            fd = (_critical_loss64(xp, edges, weights, cx.edge_i, cx.edge_j)
                  - _critical_loss64(xm, edges, weights, cx.edge_i, cx.edge_j)) / (2 * h)
        assert abs(fd.item() - ours[i, k].item()) <= 1e-3 * max(1.0, abs(fd.item())), (i, k)


def test_full_pipeline_finite_differences():
    """Recompute the diagrams at x +- h e_k: where the critical edges stay the
    same, the change of the loss matches the gradient."""
    from flash_ph import rips_persistence
    print("[gpu] autograd full-pipeline FD", flush=True)
    x0 = cases.clifford(200, 7, noise=0.02).astype(np.float64)
    t = 0.7

    def h1_total(xnp):
        d = rips_persistence(torch.from_numpy(xnp.astype(np.float32)).cuda(), t, max_dim=1)[1]
        fin = torch.isfinite(d[:, 1])
        return float((d[fin, 1].double() - d[fin, 0].double()).sum())

    x = torch.from_numpy(x0.astype(np.float32)).cuda().requires_grad_(True)
    d1 = rips_persistence(x, t, max_dim=1)[1]
    fin = torch.isfinite(d1[:, 1])
    (d1[fin, 1] - d1[fin, 0]).sum().backward()
    grad = x.grad.double().cpu().numpy()
    h = 1e-3
    rng = np.random.default_rng(3)
    checked = 0
    for _ in range(40):
        i, k = int(rng.integers(200)), int(rng.integers(4))
        xp, xm = x0.copy(), x0.copy()
        xp[i, k] += h
        xm[i, k] -= h
        fd = (h1_total(xp) - h1_total(xm)) / (2 * h)
        if abs(fd - grad[i, k]) <= 2e-3 * max(1.0, abs(fd)):
            checked += 1
    # a perturbation may reorder near-tied edges (the loss is piecewise
    # smooth); most coordinates keep the pairing
    assert checked >= 30, checked


def test_batch_backward_clifford():
    """Batched clouds (Clifford torus, d=4, t=0.6, loss = total H1
    persistence): the batch gradient equals each cloud's own gradient and
    torch autograd on the recomputed lengths."""
    from flash_ph import rips_persistence, rips_persistence_batch
    print("[gpu] batch autograd clifford", flush=True)
    clouds_np = cases.clifford_batch(6, 300, seed=100)
    xs = [torch.from_numpy(c).cuda().requires_grad_(True) for c in clouds_np]
    dg = rips_persistence_batch(xs, 0.6, max_dim=1)
    loss = 0
    for per in dg:
        d1 = per[1]
        fin = torch.isfinite(d1[:, 1])
        loss = loss + (d1[fin, 1] - d1[fin, 0]).sum()
    loss.backward()
    for c, xn in enumerate(clouds_np):
        y = torch.from_numpy(xn).cuda().requires_grad_(True)
        d1 = rips_persistence(y, 0.6, max_dim=1)[1]
        fin = torch.isfinite(d1[:, 1])
        (d1[fin, 1] - d1[fin, 0]).sum().backward()
        assert torch.equal(xs[c].grad.view(torch.int32), y.grad.view(torch.int32)), c


def test_grad_determinism_and_zero_paths():
    from flash_ph import rips_persistence
    print("[gpu] autograd determinism", flush=True)
    x0 = cases.randn(150, 3, 5)
    grads = []
    for _ in range(2):
        x = torch.from_numpy(x0).cuda().requires_grad_(True)
        d = rips_persistence(x, 1.0, max_dim=1)
        fin = torch.isfinite(d[1][:, 1])
        (d[0][torch.isfinite(d[0][:, 1]), 1].sum() + d[1][fin].sum()).backward()
        grads.append(x.grad.clone())
    assert torch.equal(grads[0], grads[1])
    # duplicate points: a zero-length edge gets the subgradient 0, no NaN
    y = np.concatenate([x0[:20], x0[:5]]).astype(np.float32)
    x = torch.from_numpy(y).cuda().requires_grad_(True)
    rips_persistence(x, 1.0, max_dim=1)[0][:, 1].nan_to_num(posinf=0).sum().backward()
    assert torch.isfinite(x.grad).all()
