"""GPU covariance route (scxpca_matrix_pca_covariance, the reference's
covariance_pca, scx-accel/src/pca/cpu.rs:1663-1854) against the oracle's
restatement on the same matrices.

Tolerances: the GPU Gram sums f32 products over each 256-row tile and f64
across tiles (~1e-7 relative per entry), the eigendecomposition is f64, so
eigenvalues (variance) agree to 1e-6 of the largest plus 1e-5 relative, the
variance ratio likewise, and the total variance to 1e-9 relative (both from
the same f64 column statistics). Eigenvectors are compared on the leading
block that a relative eigenvalue gap of at least 1 % separates (principal
angle <= 1e-4 rad), embeddings column by column on that block after sign
alignment (<= 1e-4 of the column norm), the mean to 1e-6 relative.
"""
from __future__ import annotations

import sys

import numpy as np
import pytest

import oracle
from scx_amd import synth
from tests.pca_metrics import align_signs, max_principal_angle, separated_rank
from tests.test_gpu_shapes import MATRICES

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def cfg1_head():
    x = synth.generate(synth.CFG1, n_obs=3000, workers=1)
    return x.indptr, x.indices, x.data, x.shape


def shape_case(name):
    c = MATRICES[name]()
    return c.indptr, c.indices, c.data, c.shape


CASES = {
    "cfg1_3000x2000_k50": (cfg1_head, 50, True),
    "cfg1_3000x2000_k100_uncentered": (cfg1_head, 100, False),
    "ragged_k20": (lambda: shape_case("ragged"), 20, True),
    "dense60_k20": (lambda: shape_case("dense60"), 20, True),
    "wide3k_k30": (lambda: shape_case("wide3k"), 30, True),
    "tiny_k5": (lambda: shape_case("tiny"), 5, True),
}


@pytest.mark.parametrize("case", list(CASES))
def test_covariance_route_matches_oracle(gpu, case):
    make, nc, centered = CASES[case]
    ip, ix, dv, shape = make()
    say(f"case: covariance {case} {shape} nnz={len(dv)}")
    ref = oracle.covariance_pca(oracle.CsrView(ip, ix, dv, shape), n_comps=nc, zero_center=centered,
                                target_sum=1e4, log1p=True)
    mat = gpu.Matrix(ip, ix, dv, shape, target_sum=1e4, log1p=True)
    try:
        got = mat.pca_covariance(n_comps=nc, zero_center=centered)
    finally:
        mat.close()
    rv, gv = ref["variance"], got["variance"]
    assert np.all(np.abs(gv - rv) <= 1e-6 * rv[0] + 1e-5 * rv), np.abs(gv - rv).max() / rv[0]
    assert abs(got["total_variance"] - ref["total_variance"]) <= 1e-9 * ref["total_variance"]
    rr = ref["variance_ratio"]
    assert np.all(np.abs(got["variance_ratio"] - rr) <= 1e-6 * rr[0] + 1e-5 * rr)
    j = separated_rank(np.sqrt(rv), min(nc, 10), 0.01)
    say(f"{case}: separated rank {j}, var rel err {np.abs(gv - rv).max() / rv[0]:.2e}")
    if j:
        rc = ref["components"][:j].T
        gc = got["components"][:j].T.astype(np.float64)
        assert max_principal_angle(rc, gc) <= 1e-4
        re = ref["embeddings"][:, :j]
        ge = align_signs(re, got["embeddings"][:, :j].astype(np.float64), axis=0)
        err = np.abs(ge - re).max(axis=0) / np.maximum(np.linalg.norm(re, axis=0), 1e-300)
        assert err.max() <= 1e-4, err
    if centered:
        np.testing.assert_allclose(got["mean"], ref["mean"], rtol=1e-6, atol=1e-12)
    assert got["k"] == nc


def test_covariance_rejects_too_many_components(gpu):
    ip, ix, dv, shape = shape_case("tiny")
    mat = gpu.Matrix(ip, ix, dv, shape, target_sum=1e4, log1p=True)
    try:
        with pytest.raises(ValueError, match="exceeds matrix rank bound"):
            mat.pca_covariance(n_comps=31)
    finally:
        mat.close()


def test_accel_auto_routes_to_randomized(gpu):
    """On the GPU pca(method="auto") runs the randomized route, as the
    reference's GPU dispatch does (pyscx/src/accel/pca.rs:384-389): its result
    equals method="randomized" bit for bit; an explicit method="covariance"
    writes the covariance route's result (variance equal to the direct
    covariance call), which differs from the approximate randomized one."""
    from scx_amd import accel
    from scx_amd.adata import AnnDataLite

    ip, ix, dv, shape = shape_case("ragged")
    import scipy.sparse as sp

    raw = sp.csr_matrix((dv, ix, ip), shape=shape)
    mat = gpu.Matrix(ip, ix, dv, shape, target_sum=1e4, log1p=True)
    try:
        direct = mat.pca_covariance(n_comps=10)
    finally:
        mat.close()
    results = {}
    for method in ("auto", "covariance", "randomized"):
        a = AnnDataLite(raw.copy())
        accel.normalize_total(a, target_sum=1e4)
        accel.log1p(a)
        say(f"case: accel.pca method={method}")
        accel.pca(a, n_comps=10, device="gpu", method=method)
        results[method] = a.uns["pca"]["variance"]
    np.testing.assert_array_equal(results["auto"], results["randomized"])
    np.testing.assert_allclose(results["covariance"], direct["variance"], rtol=1e-12)
    assert not np.allclose(results["randomized"], direct["variance"], rtol=1e-12, atol=0)
