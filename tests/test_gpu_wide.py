"""k' = n_comps + n_oversamples > 128: the wide-panel route (128-column
blocks through the KP = 128 SpMMs; QR and the Rayleigh-Ritz SVD on rocSOLVER)
against the oracle's f64 randomized PCA on the same Omega. The reference puts
no panel limit on k' (scx-gpu/src/gpu_pca.rs:433-446 checks only
n_components <= min(n_obs, n_vars); k at scx-accel/src/pca/cpu.rs:549).

Bars as the narrow route's (FP32 device arithmetic against f64): singular
values <= 1e-4 relative (the top 100), largest principal angle <= 1e-3 rad on
the gap-separated leading block, scores <= 1e-3 relative there, mean 1e-6.
"""
from __future__ import annotations

import sys

import numpy as np
import pytest
import scipy.sparse as sp

import oracle
from scx_amd import synth
from tests.pca_metrics import max_principal_angle, score_rel_err, separated_rank, sigma_rel_err

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def head():
    return synth.generate(synth.CFG1, n_obs=3000, workers=1)


@pytest.fixture(scope="module")
def head_mat(gpu, head):
    say("case: Matrix(cfg1 head 3000)")
    m = gpu.Matrix(head.indptr, head.indices, head.data, head.shape, target_sum=1e4, log1p=True)
    yield m
    m.close()


@pytest.fixture(scope="module")
def ref200(head):
    say("oracle rPCA k'=200 q=2")
    return oracle.randomized_pca(oracle.CsrView(head.indptr, head.indices, head.data, head.shape), n_comps=190,
                                 n_oversamples=10, n_power_iterations=2, zero_center=True, seed=0, target_sum=1e4,
                                 log1p=True)


def _check(ref, got, nc, label, sig_bar=1e-4):
    se = sigma_rel_err(ref["sigma_all"][:100], got["sigma"][:100])
    j = max(separated_rank(ref["sigma_all"], 10), 1)
    ang = max_principal_angle(ref["components"].T[:, :j], got["components"].T[:, :j].astype(np.float64))
    sc = score_rel_err(ref["embeddings"][:, :j], got["embeddings"][:, :j].astype(np.float64)).max()
    say(f"{label}: sigma(top100) {se.max():.2e}  angle(top{j}) {ang:.2e}  scores {sc:.2e}")
    assert got["k"] == len(ref["sigma_all"])
    assert got["embeddings"].shape[1] == nc and got["components"].shape[0] == nc
    assert se.max() <= sig_bar
    assert ang <= 1e-3
    assert sc <= 1e-3
    np.testing.assert_allclose(got["mean"], ref["mean"], rtol=1e-6, atol=1e-12)
    np.testing.assert_allclose(got["variance_ratio"][:100], ref["variance_ratio"][:100], rtol=2e-4)


@pytest.mark.parametrize("qr_method", [0, 1])
def test_wide_matches_oracle(head_mat, ref200, qr_method):
    say(f"case: wide k'=200 qr={qr_method}")
    got = head_mat.pca(n_comps=190, n_oversamples=10, n_power_iterations=2, seed=0, qr_method=qr_method)
    _check(ref200, got, 190, f"wide qr={qr_method}")
    # components orthonormal (the f64 SVD's left vectors, rounded to f32)
    c = got["components"].astype(np.float64)
    assert np.abs(c @ c.T - np.eye(190)).max() <= 1e-5


def test_wide_three_blocks_uncentered(gpu, head):
    """k' = 300 (three 128-column blocks, the last one partly padding), no centering."""
    say("case: wide k'=300 uncentered")
    view = oracle.CsrView(head.indptr, head.indices, head.data, head.shape)
    ref = oracle.randomized_pca(view, n_comps=280, n_oversamples=20, n_power_iterations=1, zero_center=False,
                                seed=7, target_sum=1e4, log1p=True)
    m = gpu.Matrix(head.indptr, head.indices, head.data, head.shape, target_sum=1e4, log1p=True)
    try:
        got = m.pca(n_comps=280, n_oversamples=20, n_power_iterations=1, zero_center=False, seed=7, qr_method=0)
    finally:
        m.close()
    se = sigma_rel_err(ref["sigma_all"][:100], got["sigma"][:100])
    ang = max_principal_angle(ref["components"].T[:, :5], got["components"].T[:, :5].astype(np.float64))
    say(f"wide k'=300: sigma(top100) {se.max():.2e}  angle(top5) {ang:.2e}")
    assert got["k"] == 300 and se.max() <= 1e-4 and ang <= 1e-3
    assert got["mean"] is None


def test_wide_bf16_sketch(head_mat):
    """The BF16 sketch form on wide panels against the FP32 form."""
    say("case: wide k'=200 bf16 sketch")
    f32 = head_mat.pca(n_comps=190, n_oversamples=10, n_power_iterations=2, seed=0, qr_method=0)
    b16 = head_mat.pca(n_comps=190, n_oversamples=10, n_power_iterations=2, seed=0, qr_method=0,
                       precision=1)
    se = sigma_rel_err(f32["sigma"][:50], b16["sigma"][:50]).max()
    say(f"wide bf16 vs fp32: sigma(top50) {se:.2e}")
    assert se <= 1e-4


def test_wide_streamed_matches_resident(gpu, head, head_mat):
    ip, ix, dv = head.indptr.astype(np.int64), head.indices.astype(np.uint32), head.data.astype(np.float32)
    kw = dict(n_comps=150, n_oversamples=10, n_power_iterations=2, seed=0, qr_method=1)
    say("case: wide k'=160 streamed shard_rows=700")
    got = gpu.randomized_pca_streamed(ip, ix, dv, head.shape, 700, target_sum=1e4, log1p=True, **kw)
    res = head_mat.pca(**kw)
    se = sigma_rel_err(res["sigma"][:100], got["sigma"][:100]).max()
    say(f"wide streamed vs resident: sigma(top100) {se:.2e}")
    assert se <= 1e-5
    np.testing.assert_allclose(got["mean"], res["mean"], rtol=1e-12, atol=1e-300)


def test_wide_covariance_route(gpu, head):
    """method='covariance' with 150 components: the embeddings in two
    128-component slices."""
    say("case: covariance route n_comps=150")
    ref = oracle.covariance_pca(oracle.CsrView(head.indptr, head.indices, head.data, head.shape), n_comps=150,
                                target_sum=1e4, log1p=True)
    m = gpu.Matrix(head.indptr, head.indices, head.data, head.shape, target_sum=1e4, log1p=True)
    try:
        got = m.pca_covariance(n_comps=150)
    finally:
        m.close()
    np.testing.assert_allclose(got["variance"], ref["variance"], rtol=1e-5, atol=1e-6 * ref["variance"][0])
    j = max(separated_rank(np.sqrt(ref["variance"]), 10), 1)
    ang = max_principal_angle(ref["components"].T[:, :j], got["components"].T[:, :j].astype(np.float64))
    say(f"covariance 150: angle(top{j}) {ang:.2e}")
    assert ang <= 1e-4
    # embeddings of the last slice: (X~ - mu) v_j, sign-aligned with the oracle's
    e = score_rel_err(ref["embeddings"][:, 128:150], got["embeddings"][:, 128:150].astype(np.float64))
    say(f"covariance 150: slice-2 embeddings vs (X~ - mu) V_ref {e.max():.2e}")
    # individual tail eigenvectors can rotate inside near-degenerate pairs; the
    # embedding norms (sqrt((n-1) lambda)) are rotation-free
    np.testing.assert_allclose(np.linalg.norm(got["embeddings"][:, 128:150].astype(np.float64), axis=0),
                               np.linalg.norm(ref["embeddings"][:, 128:150], axis=0), rtol=1e-3)


def test_wide_cholesky_non_spd(gpu):
    """k' above the rank (10 distinct cells repeated 30 times, 200 genes):
    CholeskyQR2 reports the non-SPD error on wide panels too; the Householder
    mode returns the rank-10 spectrum with zeros beyond it."""
    rng = np.random.default_rng(5)
    base = rng.negative_binomial(2, 0.4, size=(10, 200)).astype(np.float32)
    x = sp.csr_matrix(np.repeat(base, 30, axis=0))
    m = gpu.Matrix(x.indptr.astype(np.int64), x.indices.astype(np.uint32), x.data, x.shape, target_sum=1e4,
                   log1p=True)
    try:
        say("case: wide rank-deficient cholesky")
        with pytest.raises(gpu.LinAlgError, match="CholeskyQR2 failed: non-SPD"):
            m.pca(n_comps=140, n_oversamples=10, n_power_iterations=2, seed=0, qr_method=1)
        say("case: wide rank-deficient householder")
        got = m.pca(n_comps=140, n_oversamples=10, n_power_iterations=2, seed=0, qr_method=0)
    finally:
        m.close()
    s = got["sigma"]
    assert np.all(np.isfinite(got["embeddings"])) and np.all(np.isfinite(got["components"]))
    assert np.all(s[9:] <= 1e-4 * s[0]) and s[8] > 1e-3 * s[0]


def test_accel_pca_n_comps_150(gpu, head):
    """pyscx.accel.pca(n_comps=150) on the GPU: the drop-in call the panel
    limit used to refuse."""
    from scx_amd import accel
    from scx_amd.adata import AnnDataLite

    a = AnnDataLite(sp.csr_matrix((head.data, head.indices.astype(np.int64), head.indptr), shape=head.shape))
    accel.normalize_total(a, target_sum=1e4)
    accel.log1p(a)
    say("case: accel.pca n_comps=150")
    accel.pca(a, n_comps=150, device="gpu")
    assert a.obsm["X_pca"].shape == (3000, 150) and a.varm["PCs"].shape == (2000, 150)
    v = a.uns["pca"]["variance"]
    assert np.all(np.diff(v) <= 1e-9 * v[0]) and a.uns["pca"]["variance_ratio"].sum() <= 1.0
