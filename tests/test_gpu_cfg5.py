"""Full-transcriptome shape: the first 20,000 cells of the 2,000,000 x 33,538
workload generator (seed 1, ~6% density), k=100, p=20 (k' = 120, the KP=128
kernels), 4 power iterations, through the wide-gene paths (row-parallel
forward SpMM, general tile-CSC build). FP32 and BF16-sketch forms against the
oracle's f64 rPCA on the same matrix and Omega: singular values <= 1e-4
relative (all 100) for both forms; FP32: top-10 principal angle <= 1e-3 rad
and top-10 scores <= 1e-3 relative after sign alignment. The BF16 sketch
perturbs the subspace the power iterations find, so its angle is checked
(<= 1e-3 rad) on the leading subspace that a relative singular-value gap of at
least 2% separates: at 20,000 cells this workload has one dominant component
and then a flat bulk (gaps < 0.5%), where any top-j subspace inside the bulk
is not determined by the data. Column means to 1e-6.
"""
from __future__ import annotations

import os
import sys

import numpy as np
import pytest

import oracle
from scx_amd import synth
from tests.pca_metrics import max_principal_angle, score_rel_err, separated_rank, sigma_rel_err

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def sub(gpu):
    w = synth.CFG5
    say("case: generate 20k rows of the 2M x 33,538 workload")
    x = synth.generate(w, n_obs=20000)
    nthreads = max(1, min(16, len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else 1))
    say("case: oracle rPCA 20k x 33,538 k=100 p=20 q=4")
    ref = oracle.randomized_pca(oracle.CsrView(x.indptr, x.indices, x.data, x.shape), n_comps=w.n_comps,
                                n_oversamples=w.n_oversamples, n_power_iterations=w.n_power_iterations, seed=0,
                                target_sum=w.target_sum, log1p=True, nthreads=nthreads)
    m = gpu.Matrix(x.indptr, x.indices, x.data, x.shape, target_sum=w.target_sum, log1p=True)
    yield w, x, m, ref
    m.close()


@pytest.mark.parametrize("precision", [0, 1])
def test_cfg5_subsample_pca(sub, precision):
    w, x, m, ref = sub
    say(f"case: cfg5 subsample rPCA precision={precision}")
    got = m.pca(n_comps=w.n_comps, n_oversamples=w.n_oversamples, n_power_iterations=w.n_power_iterations, seed=0,
                qr_method=1, precision=precision)
    se = sigma_rel_err(ref["sigma_all"][: w.n_comps], got["sigma"][: w.n_comps])
    ang = max_principal_angle(ref["components"].T[:, :10], got["components"].T[:, :10].astype(np.float64))
    j = 10 if precision == 0 else separated_rank(ref["sigma_all"], 10, 0.02)
    ang_j = max_principal_angle(ref["components"].T[:, :j], got["components"].T[:, :j].astype(np.float64)) if j else 0.0
    say(f"precision={precision}: sigma rel err max {se.max():.2e}, angle top10 {ang:.2e}, "
        f"angle top{j} {ang_j:.2e}")
    assert se.max() <= 1e-4
    assert ang_j <= 1e-3
    np.testing.assert_allclose(got["mean"], ref["mean"], rtol=1e-6, atol=1e-12)
    if precision == 0:
        sc = score_rel_err(ref["embeddings"][:, :10], got["embeddings"][:, :10].astype(np.float64))
        say(f"score rel err top10 {sc.max():.2e}")
        assert sc.max() <= 1e-3


@pytest.mark.skipif(os.environ.get("SCX_FULL_CFG5") != "1",
                    reason="full 2,000,000 x 33,538 matrix (4.0e9 nonzeros, 32 GB of host CSR): "
                           "set SCX_FULL_CFG5=1 to run")
def test_cfg5_full_size_identities(gpu):
    """Size-independent checks at the full size, through device pointers:
    adjoint <X~^T Q, M> = <Q, X~ M> with positive Q and M (both whole-matrix
    passes, f64 host sums, 2e-5 relative); the gene sum of the transposed pass, sum_c (X~^T Q)[c, :] =
    (X~ 1)^T Q with X~ 1 from the oracle's transform (1e-5 relative); and the
    FP32 and BF16-sketch rPCA (k=100, p=20, q=4) agreeing on the singular values
    (1e-4 relative) and the top-10 subspace of the leading separated block."""
    import torch

    w = synth.CFG5
    say("case: generate the full 2M x 33,538 workload")
    x = synth.generate(w)
    say(f"nnz={x.nnz:,}")
    dev = torch.device("cuda:0")
    t = (torch.from_numpy(x.indptr).to(dev), torch.from_numpy(x.indices.view(np.int32)).to(dev),
         torch.from_numpy(x.data).to(dev))
    torch.cuda.synchronize()
    m = gpu.Matrix(*t, x.shape, target_sum=w.target_sum, log1p=True, device=0)
    try:
        n, nv = x.shape
        rng = np.random.default_rng(3)
        k = 16
        # Positive panels for the adjoint: with ~1.2e5 rows per gene, f32
        # rounding of a sign-mixed sum leaves ~1e-5 of the (cancelling) total,
        # while positive sums keep it near 5e-6 relative, so a lost tile
        # (1.3e-4 of the rows) still shows at 2e-5.
        q = rng.uniform(0.0, 1.0, (n, k)) / n
        panel = rng.uniform(0.0, 1.0, (nv, k))
        say("case: full transpose + forward, uncentered")
        z = m.spmm(q, transpose=True)
        y = m.spmm(panel, transpose=False)
        a, b = float(np.sum(z * panel)), float(np.sum(q * y))
        say(f"adjoint: {a!r} vs {b!r}")
        assert abs(a - b) <= 2e-5 * max(abs(a), abs(b))
        qs = rng.standard_normal((n, k)) / np.sqrt(n)
        zs = m.spmm(qs, transpose=True)
        view = oracle.CsrView(x.indptr, x.indices, x.data, x.shape)
        xrow = oracle.row_sums(view.with_data(oracle.transform(view, w.target_sum, True)))
        ref = xrow @ qs
        assert np.abs(zs.sum(axis=0) - ref).max() <= 1e-5 * np.abs(ref).max()
        del z, y, q, zs, qs
        say("case: full rPCA FP32 and BF16 sketch")
        kw = dict(n_comps=w.n_comps, n_oversamples=w.n_oversamples, n_power_iterations=w.n_power_iterations,
                  seed=0, qr_method=1)
        f32 = m.pca(precision=0, **kw)
        b16 = m.pca(precision=1, **kw)
        se = sigma_rel_err(f32["sigma"][: w.n_comps], b16["sigma"][: w.n_comps])
        j = separated_rank(f32["sigma"], 10, 0.02)
        ang = max_principal_angle(f32["components"].T[:, :j].astype(np.float64),
                                  b16["components"].T[:, :j].astype(np.float64)) if j else 0.0
        say(f"FP32 vs BF16: sigma rel max {se.max():.2e}, top{j} angle {ang:.2e}, "
            f"device s {f32['seconds_device']:.3f} / {b16['seconds_device']:.3f}")
        assert se.max() <= 1e-4
        assert ang <= 1e-3
    finally:
        m.close()


@pytest.mark.skipif(os.environ.get("SCX_FULL_CFG5") != "1",
                    reason="full 2,000,000 x 33,538 matrix (4.0e9 nonzeros, 32 GB of host CSR): "
                           "set SCX_FULL_CFG5=1 to run")
def test_cfg5_full_size_streamed_matches_resident(gpu):
    """The streamed operator at the full size (8 shards of 250,000 rows, the
    whole CSR over PCIe on each of the 11 passes) against the resident
    matrix: singular values to 1e-5, the gap-separated leading block to
    1e-4 rad (measured: sigma 2.5e-7; top-3 8.4e-5 rad with sigma 2 and 3
    0.04 % apart), mean to 1e-12."""
    w = synth.WORKLOADS["pca_2m_x_33538_k100"]
    x = synth.generate(w)
    say(f"case: full-size streamed vs resident {x.shape} nnz={x.nnz}")
    kw = dict(n_comps=w.n_comps, n_oversamples=w.n_oversamples, n_power_iterations=w.n_power_iterations, seed=0,
              qr_method=1)
    m = gpu.Matrix(x.indptr, x.indices, x.data, x.shape, target_sum=w.target_sum, log1p=True)
    try:
        res = m.pca(**kw)
    finally:
        m.close()
    got = gpu.randomized_pca_streamed(x.indptr, x.indices, x.data, x.shape, 250_000, target_sum=w.target_sum,
                                      log1p=True, **kw)
    se = sigma_rel_err(res["sigma"][:10], got["sigma"][:10]).max()
    # the leading block a 1 % gap separates (sigma 2 and 3 nearly coincide here)
    j = max(1, separated_rank(res["sigma"], 10, 0.01))
    ang = max_principal_angle(res["components"].T[:, :j].astype(np.float64), got["components"].T[:, :j].astype(np.float64))
    say(f"full-size streamed: sigma vs resident {se:.2e}, angle top{j} {ang:.2e}, "
        f"device s {res['seconds_device']:.3f} / {got['seconds_device']:.3f}")
    assert se <= 1e-5
    assert ang <= 1e-4
    np.testing.assert_allclose(got["mean"], res["mean"], rtol=1e-12, atol=1e-300)
