"""The committed golden vectors (tests/golden/golden_v1.npz, written by
tests/golden/make_golden.py from the oracle).

CPU: the oracle reproduces every vector bit for bit, and each vector is
checked against an independent restatement:
* X~ of test_csr_10x5 (scx-accel/src/pca/cpu.rs:2326): numpy's float32 log1p
  of f32(f64(x) * 1e4 / row_sum), to 1 ulp;
* the Omega prefixes: first two moments of 4,096 standard normal draws;
* the SpMM outputs: dense float64 numpy products of the same X~ (1e-12);
* the PCA: the input regenerates to the stored SHA-256; every Rayleigh-Ritz
  singular value is at most the exact one of the dense centered X~ (the
  projection B = X~^T Q onto an orthonormal basis cannot increase them) and
  the top 10 lie within 1e-2 of them; variance = sigma^2 / (n - 1); the
  variance ratios sum to at most 1.
GPU (through the C ABI): device Omega bit for bit; SpMM to 2e-5 of the f64
magnitude; PCA singular values <= 1e-4 relative, top-10 principal angle
<= 1e-3 rad, mean to 1e-6.
"""
from __future__ import annotations

import os
import sys

import numpy as np
import pytest
import scipy.sparse as sp

import oracle
from tests.pca_metrics import max_principal_angle, sigma_rel_err

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(HERE, "golden"))
import make_golden  # noqa: E402

G = np.load(os.path.join(HERE, "golden", "golden_v1.npz"), allow_pickle=False)


def _xt_dense(indptr, indices, data, shape):
    v = oracle.CsrView(indptr, indices, data, shape)
    xt = oracle.transform(v, 1e4, True).astype(np.float64)
    return sp.csr_matrix((xt, indices, indptr), shape=shape).toarray()


def test_xt_10x5():
    c = make_golden.CSR_10X5
    v = oracle.CsrView(c["indptr"], c["indices"], c["data"], c["shape"])
    got = oracle.transform(v, 1e4, True).astype(np.float32)
    assert np.array_equal(got.view(np.uint32), G["xt_10x5"].view(np.uint32))
    rows = np.repeat(np.arange(10), np.diff(c["indptr"]))
    s = np.add.reduceat(c["data"].astype(np.float64), c["indptr"][:-1])
    scaled = (c["data"].astype(np.float64) * (1e4 / s)[rows]).astype(np.float32)
    ref = np.log1p(scaled)
    ulp = np.abs(G["xt_10x5"].view(np.int32).astype(np.int64) - ref.view(np.int32).astype(np.int64))
    assert ulp.max() <= 1


@pytest.mark.parametrize("seed", make_golden.OMEGA_SEEDS)
def test_omega_prefix(seed):
    got = oracle.random_gaussian(64, 64, seed).reshape(-1)
    want = G[f"omega_s{seed}"]
    assert np.array_equal(got.view(np.uint64), want.view(np.uint64))
    assert abs(want.mean()) < 0.06 and abs(want.std() - 1.0) < 0.05


def _spmm_case():
    shape = (200, 50)
    ip, ix, dv = G["spmm_indptr"], G["spmm_indices"], G["spmm_data"]
    return ip, ix, dv, shape


def test_spmm_oracle_and_dense():
    ip, ix, dv, shape = _spmm_case()
    v = oracle.CsrView(ip, ix, dv, shape)
    xt = v.with_data(oracle.transform(v, 1e4, True))
    m, q, mu = G["spmm_m"], G["spmm_q"], G["spmm_mu"]
    for name, got in (("spmm_fwd", oracle.spmm_forward(xt, m)), ("spmm_fwd_centered", oracle.spmm_forward(xt, m, mu)),
                      ("spmm_t", oracle.spmm_transpose(xt, q)),
                      ("spmm_t_centered", oracle.spmm_transpose(xt, q, mu))):
        assert np.array_equal(got, G[name]), name
    d = _xt_dense(ip, ix, dv, shape)
    assert np.allclose(mu, d.mean(axis=0), rtol=1e-13, atol=1e-15)
    for name, ref in (("spmm_fwd", d @ m), ("spmm_fwd_centered", (d - mu) @ m), ("spmm_t", d.T @ q),
                      ("spmm_t_centered", (d - mu).T @ q)):
        assert np.abs(G[name] - ref).max() <= 1e-12 * np.abs(ref).max(), name


@pytest.fixture(scope="module")
def pca_in():
    ip, ix, dv, shp = make_golden.pca_input()
    assert make_golden.digest(ip, ix, dv) == str(G["pca_input_sha256"]), "workload generator changed"
    return ip, ix, dv, shp


def test_pca_oracle_and_bounds(pca_in):
    ip, ix, dv, shp = pca_in
    res = oracle.randomized_pca(oracle.CsrView(ip, ix, dv, shp), n_comps=50, n_oversamples=10,
                                n_power_iterations=2, seed=0, target_sum=1e4, log1p=True)
    assert np.array_equal(res["sigma_all"], G["pca_sigma"])
    assert np.array_equal(res["mean"], G["pca_mean"])
    assert np.array_equal(res["components"][:10], G["pca_components10"])
    assert np.array_equal(res["embeddings"][:, :10], G["pca_embeddings10"])
    d = _xt_dense(ip, ix, dv, shp)
    exact = np.linalg.svd(d - d.mean(axis=0), compute_uv=False)
    sig = G["pca_sigma"]
    assert np.all(sig <= exact[: sig.size] * (1 + 1e-12))
    assert np.all(np.abs(sig[:10] - exact[:10]) <= 1e-2 * exact[:10])
    n = shp[0]
    assert np.allclose(G["pca_variance"], sig[:50] ** 2 / (n - 1), rtol=1e-12)
    assert G["pca_variance_ratio"].sum() <= 1.0 + 1e-12


@pytest.mark.gpu
@pytest.mark.parametrize("seed", make_golden.OMEGA_SEEDS)
def test_gpu_omega_prefix(gpu, seed):
    from scx_amd import _lib

    got = _lib.gaussian_device(64, 64, seed, device=0).reshape(-1)
    assert np.array_equal(got.view(np.uint64), G[f"omega_s{seed}"].view(np.uint64))


@pytest.mark.gpu
def test_gpu_spmm(gpu):
    ip, ix, dv, shape = _spmm_case()
    mat = gpu.Matrix(ip, ix, dv, shape, target_sum=1e4, log1p=True)
    try:
        mu = G["spmm_mu"]
        for name, kw in (("spmm_fwd", dict(panel=G["spmm_m"], transpose=False)),
                         ("spmm_fwd_centered", dict(panel=G["spmm_m"], transpose=False, means=mu)),
                         ("spmm_t", dict(panel=G["spmm_q"], transpose=True)),
                         ("spmm_t_centered", dict(panel=G["spmm_q"], transpose=True, means=mu))):
            got = mat.spmm(kw.pop("panel"), **kw)
            ref = G[name]
            assert np.abs(got - ref).max() <= 2e-5 * np.abs(ref).max(), name
    finally:
        mat.close()


@pytest.mark.gpu
def test_gpu_pca(gpu, pca_in):
    ip, ix, dv, shp = pca_in
    mat = gpu.Matrix(ip, ix, dv, shp, target_sum=1e4, log1p=True)
    try:
        got = mat.pca(n_comps=50, n_oversamples=10, n_power_iterations=2, seed=0, qr_method=1, precision=0)
    finally:
        mat.close()
    assert sigma_rel_err(G["pca_sigma"][:50], got["sigma"][:50]).max() <= 1e-4
    assert max_principal_angle(G["pca_components10"].T, got["components"][:10].T.astype(np.float64)) <= 1e-3
    np.testing.assert_allclose(got["mean"], G["pca_mean"], rtol=1e-6, atol=1e-12)
