"""The differentiable projection Z = (X~ - 1 mu^T) W (scx_amd.projection).

* 64 cells x 200 genes slice of the 10k x 2k workload generator: the HIP
  forward and hand-written backward against (a) torch.sparse in float64 on the
  CPU with autograd (reference gradient) and (b) central finite differences of
  the float64 reference function (step 1e-3; the function is linear in W and
  mu, so the difference quotient has no truncation error). Tolerance 1e-5 of
  the largest magnitude (FP32 kernels against f64).
* one 4,096-cell x 2,000-gene minibatch (generator seed = batch id): Z, dW and
  dmu against scipy float64, tolerance 2e-5 of the largest magnitude.
"""
from __future__ import annotations

import sys

import numpy as np
import pytest
import scipy.sparse as sp

import oracle
from scx_amd import synth

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _slice(csr, n_rows, n_cols):
    m = csr.to_scipy()[:n_rows, :n_cols].tocsr()
    m.sort_indices()
    return m


def _xtilde(m):
    v = oracle.CsrView(m.indptr.astype(np.int64), m.indices.astype(np.uint32), m.data.astype(np.float32), m.shape)
    return sp.csr_matrix((oracle.transform(v, 1e4, True).astype(np.float64), m.indices, m.indptr), shape=m.shape)


def _device_matrix(gpu, m):
    import torch

    dev = torch.device("cuda:0")
    t = (torch.from_numpy(m.indptr.astype(np.int64)).to(dev),
         torch.from_numpy(m.indices.astype(np.int32)).to(dev),
         torch.from_numpy(m.data.astype(np.float32)).to(dev))
    torch.cuda.synchronize()
    return gpu.Matrix(*t, m.shape, target_sum=1e4, log1p=True, device=0), t


def _run(gpu, m, w_np, mu_np, dz_np):
    import torch

    from scx_amd.projection import project

    dev = torch.device("cuda:0")
    mat, keep = _device_matrix(gpu, m)
    w = torch.from_numpy(w_np.astype(np.float32)).to(dev).requires_grad_(True)
    mu = torch.from_numpy(mu_np.astype(np.float32)).to(dev).requires_grad_(True)
    z = project(mat, w, mu)
    z.backward(torch.from_numpy(dz_np.astype(np.float32)).to(dev))
    torch.cuda.synchronize()
    out = z.detach().cpu().double().numpy(), w.grad.cpu().double().numpy(), mu.grad.cpu().double().numpy()
    mat.close()
    return out


def _close(got, ref, tol):
    return np.abs(got - ref).max() <= tol * np.abs(ref).max()


def test_slice_against_torch_sparse_and_finite_differences(gpu):
    import torch

    base = synth.generate(synth.CFG1, n_obs=64, workers=1)
    m = _slice(base, 64, 200)
    rng = np.random.default_rng(4)
    k = 64
    w_np = rng.standard_normal((200, k)) / np.sqrt(200)
    xt = _xtilde(m)
    mu_np = np.asarray(xt.mean(axis=0)).ravel().astype(np.float32).astype(np.float64)
    w_np = w_np.astype(np.float32).astype(np.float64)
    dz_np = rng.standard_normal((64, k)).astype(np.float32).astype(np.float64)
    say("case: projection slice 64 x 200, k=64")
    z, dw, dmu = _run(gpu, m, w_np, mu_np, dz_np)

    # (a) torch.sparse float64 reference with autograd, on the CPU
    xs = torch.sparse_csr_tensor(torch.from_numpy(xt.indptr.astype(np.int64)),
                                 torch.from_numpy(xt.indices.astype(np.int64)),
                                 torch.from_numpy(xt.data), size=xt.shape, dtype=torch.float64)
    w64 = torch.from_numpy(w_np).requires_grad_(True)
    mu64 = torch.from_numpy(mu_np).requires_grad_(True)
    zr = torch.sparse.mm(xs, w64) - (mu64 @ w64)[None, :]
    zr.backward(torch.from_numpy(dz_np))
    assert _close(z, zr.detach().numpy(), 1e-5)
    assert _close(dw, w64.grad.numpy(), 1e-5)
    assert _close(dmu, mu64.grad.numpy(), 1e-5)

    # (b) central finite differences of L = <dZ, Z(W, mu)> in float64
    xd = xt.toarray()

    def loss(w_, mu_):
        return float(np.sum(dz_np * ((xd - mu_[None, :]) @ w_)))

    h = 1e-3
    for (i, j) in zip(rng.integers(0, 200, 24), rng.integers(0, k, 24)):
        e = np.zeros_like(w_np)
        e[i, j] = h
        fd = (loss(w_np + e, mu_np) - loss(w_np - e, mu_np)) / (2 * h)
        assert abs(dw[i, j] - fd) <= 1e-5 * np.abs(dw).max(), (i, j, dw[i, j], fd)
    for i in rng.integers(0, 200, 12):
        e = np.zeros_like(mu_np)
        e[i] = h
        fd = (loss(w_np, mu_np + e) - loss(w_np, mu_np - e)) / (2 * h)
        assert abs(dmu[i] - fd) <= 1e-5 * np.abs(dmu).max(), (i, dmu[i], fd)


@pytest.mark.parametrize("batch,k", [(3, 64), (7, 100)])
def test_minibatch_4096(gpu, batch, k):
    m = synth.generate(synth.CFG1, n_obs=4096, seed=batch, workers=1).to_scipy()
    m.sort_indices()
    rng = np.random.default_rng(batch)
    w_np = (rng.standard_normal((m.shape[1], k)) / np.sqrt(m.shape[1])).astype(np.float32).astype(np.float64)
    xt = _xtilde(m)
    mu_np = np.asarray(xt.mean(axis=0)).ravel().astype(np.float32).astype(np.float64)
    dz_np = rng.standard_normal((m.shape[0], k)).astype(np.float32).astype(np.float64)
    say(f"case: projection batch seed={batch} 4096 x 2000 k={k}")
    z, dw, dmu = _run(gpu, m, w_np, mu_np, dz_np)
    zr = xt @ w_np - (mu_np @ w_np)[None, :]
    dwr = xt.T @ dz_np - np.outer(mu_np, dz_np.sum(axis=0))
    dmur = -(w_np @ dz_np.sum(axis=0))
    assert _close(z, zr, 2e-5)
    assert _close(dw, dwr, 2e-5)
    assert _close(dmu, dmur, 2e-5)
