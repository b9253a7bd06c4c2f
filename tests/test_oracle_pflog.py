"""The PFlog (v4) oracle pinned on the CPU: the reference's own known answers
(scx-accel/src/pflog_tests.rs, pflog_reference_tests.rs, pca/cpu.rs:2726-2830),
its simulated-count fixtures with a known generating alpha
(tests/golden/pflog_alpha_fixtures.npz), and dense brute force (the
materialised Z = delta + b 1^T, its SVD and its variance)."""
from __future__ import annotations

import os

import numpy as np
import pytest
import scipy.sparse as sp

import oracle

GOLDEN = os.path.join(os.path.dirname(__file__), "golden", "pflog_alpha_fixtures.npz")

FIXTURE = np.array([[0, 1, 3, 0], [2, 0, 0, 5], [1, 1, 1, 1]], dtype=np.float32)          # pflog_tests.rs:87-93
PCA_FIXTURE = np.array([[0, 1, 3, 0], [2, 0, 0, 5], [1, 1, 1, 1], [4, 2, 0, 1], [0, 0, 6, 2], [3, 3, 1, 0]],
                       dtype=np.float32)                                                  # pflog_tests.rs:322-331


def dense_v4(rows: np.ndarray, four_alpha: float) -> np.ndarray:
    """z_ij = log1p(4a x_ij) - mean_k log1p(4a x_ik), f64 (pflog_tests.rs:47-59)."""
    logs = np.log1p(four_alpha * rows.astype(np.float64))
    return logs - logs.mean(axis=1, keepdims=True)


def reconstruct(x: sp.csr_matrix, four_alpha: float, form: int) -> np.ndarray:
    delta = oracle.pflog_delta(x, four_alpha, form)
    d = sp.csr_matrix((delta, x.indices, x.indptr), shape=x.shape)
    base = oracle.pflog_baseline(d)
    return d.toarray().astype(np.float64) + base[:, None], base


@pytest.mark.parametrize("form", [0, 1])
@pytest.mark.parametrize("four_alpha", [0.4, 2.0, 4.0, 8.0])
def test_delta_plus_baseline_is_the_dense_transform(form, four_alpha):
    # dense_equivalence_default_alpha / _various_alpha (pflog_tests.rs:98-136)
    for rows in (FIXTURE, np.array([[0, 4, 0, 2], [10, 0, 1, 0]], dtype=np.float32)):
        z, _ = reconstruct(sp.csr_matrix(rows), four_alpha, form)
        np.testing.assert_allclose(z, dense_v4(rows, four_alpha), rtol=0, atol=1e-5)
        # row_sums_are_zero (:138-153)
        assert np.abs(z.sum(axis=1)).max() <= 1e-5


def test_delta_forms_round_the_same_value_differently():
    x = sp.csr_matrix(PCA_FIXTURE)
    lazy = oracle.pflog_delta(x, 0.7, 0)
    inmem = oracle.pflog_delta(x, 0.7, 1)
    exact = np.log1p(0.7 * x.data.astype(np.float64))
    # form 1 rounds the f64 log once; form 0 takes an f32 log of an f32 product
    np.testing.assert_array_equal(inmem, exact.astype(np.float32))
    # form 0: a faithful f32 log1p (libm log1pf, as Rust's f32::ln_1p) of the f32-rounded product
    prod32 = (x.data.astype(np.float64) * 0.7).astype(np.float32).astype(np.float64)
    assert np.all(np.abs(lazy.astype(np.float64) - np.log1p(prod32)) <= np.spacing(lazy))
    assert np.abs(lazy.astype(np.float64) - exact).max() <= 4 * np.spacing(np.float32(np.log1p(4.2)))


def test_baseline_denominator_is_n_vars_and_empty_cells_are_zero():
    # centering_denominator_is_n_vars_not_nnz, empty_cell_yields_zero_baseline (:155-243)
    x = sp.csr_matrix(np.array([[0, 1, 3, 0], [0, 0, 0, 0], [1, 2, 0, 3]], dtype=np.float32))
    _, base = reconstruct(x, 4.0, 1)
    d1, d2, d3 = (np.float64(np.float32(np.log1p(4.0 * v))) for v in (1.0, 2.0, 3.0))
    assert abs(base[0] - (-(d1 + d3) / 4.0)) <= 1e-9
    assert abs(base[0] - (-(d1 + d3) / 2.0)) > 1e-6
    assert base[1] == 0.0
    assert abs(base[2] - (-(d1 + d2 + d3) / 4.0)) <= 1e-9


@pytest.mark.parametrize("zero_center,four_alpha,seed", [(True, 4.0, 42), (False, 4.0, 7), (True, 2.0, 11)])
def test_full_rank_pca_reconstructs_the_transform(zero_center, four_alpha, seed):
    # pca_zero_centered_reconstructs_centered_transform / pca_uncentered_reconstructs_transform /
    # pca_centered_reconstructs_with_general_alpha (pflog_tests.rs:336-454)
    x = sp.csr_matrix(PCA_FIXTURE)
    r = oracle.pflog_pca(x, n_comps=4, n_oversamples=0, n_power_iterations=4, zero_center=zero_center, seed=seed,
                         alpha=four_alpha / 4.0, form=0)
    z = dense_v4(PCA_FIXTURE, four_alpha)
    expected = z - z.mean(axis=0) if zero_center else z
    np.testing.assert_allclose(r["embeddings"] @ r["components"], expected, rtol=0, atol=1e-4)


@pytest.mark.parametrize("zero_center", [True, False])
def test_pflog_pca_matches_the_dense_svd(zero_center):
    rng = np.random.default_rng(3)
    dense = rng.negative_binomial(2, 0.3, size=(300, 24)).astype(np.float32)
    dense[rng.random(dense.shape) < 0.6] = 0
    dense[5] = 0  # an empty cell
    x = sp.csr_matrix(dense)
    r = oracle.pflog_pca(x, n_comps=6, n_oversamples=18, n_power_iterations=4, zero_center=zero_center, seed=1,
                         alpha=0.3, form=1)
    z, base = reconstruct(x, 1.2, 1)
    np.testing.assert_array_equal(r["baseline"], base)
    zc = z - z.mean(axis=0) if zero_center else z
    u, s, vt = np.linalg.svd(zc, full_matrices=False)
    n = z.shape[0]
    np.testing.assert_allclose(r["sigma_all"], s, rtol=1e-9, atol=1e-9)
    np.testing.assert_allclose(r["variance"], s[:6] ** 2 / (n - 1), rtol=1e-9)
    tv = (zc ** 2).sum() / (n - 1)
    assert abs(r["total_variance"] - tv) <= 1e-9 * tv
    np.testing.assert_allclose(r["variance_ratio"], s[:6] ** 2 / (n - 1) / tv, rtol=1e-9)
    for j in range(6):  # up to sign
        c = r["components"][j]
        sgn = np.sign(c @ vt[j])
        np.testing.assert_allclose(sgn * c, vt[j], atol=1e-8)
        np.testing.assert_allclose(sgn * r["embeddings"][:, j], u[:, j] * s[j], atol=1e-7)
    if zero_center:
        np.testing.assert_allclose(r["mean"], z.mean(axis=0), rtol=0, atol=1e-12)


def test_total_variance_closed_form_matches_dense():
    # pflog_total_variance_matches_dense (cpu.rs:2771-2830)
    delta = np.array([[0.10, -0.30, 0.50, 0.20], [-0.40, 0.10, 0.00, 0.60], [0.25, 0.25, -0.15, -0.05],
                      [0.70, -0.20, 0.10, -0.30], [-0.10, 0.40, -0.50, 0.30]])
    n, d = delta.shape
    base = -delta.sum(axis=1) / d
    css = (delta ** 2).sum(axis=0)
    cm = delta.mean(axis=0)
    z = delta + base[:, None]
    got_u = oracle.pflog_total_variance(css, None, base, n, d)
    assert abs(got_u - (z ** 2).sum() / (n - 1)) < 1e-9
    got_c = oracle.pflog_total_variance(css, cm, base, n, d)
    assert abs(got_c - ((z - z.mean(axis=0)) ** 2).sum() / (n - 1)) < 1e-9
    # pflog_total_variance_reports_cancellation_instead_of_a_silent_zero (cpu.rs:2726-2768)
    assert oracle.pflog_total_variance(np.full(4, 5.0), np.full(4, -1.0), np.ones(5), 5, 4) == 0.0
    assert abs(oracle.pflog_total_variance(np.full(4, 100.0), np.full(4, -1.0), np.ones(5), 5, 4) - 95.0) < 1e-12


# ------------------------------------------------------------------ alpha
def test_alpha_known_answers():
    # estimate_alpha_known_single_gene / _pools_every_gene_with_a_mean /
    # _falls_back_on_all_zero / _falls_back_when_the_pooled_median_is_not_positive
    # (pflog_tests.rs:456-543)
    e = oracle.estimate_alpha(sp.csr_matrix(np.array([[1.0], [5.0]], dtype=np.float32)))
    assert e["n_genes_used"] == 1 and not e["fell_back"]
    assert abs(e["alpha"] - 5.0 / 9.0) <= 1e-12 and abs(e["pseudocount"] - 0.45) <= 1e-12
    e = oracle.estimate_alpha(sp.csr_matrix(np.array([[1, 4, 0, 2], [5, 4, 0, 0], [3, 4, 0, 7]], dtype=np.float32)))
    assert e["n_genes_used"] == 3 and not e["fell_back"] and abs(e["alpha"] - 1.0 / 9.0) <= 1e-12
    e = oracle.estimate_alpha(sp.csr_matrix(np.zeros((2, 3), dtype=np.float32)))
    assert e["fell_back"] and e["n_genes_used"] == 0 and e["alpha"] == 0.25 and abs(e["pseudocount"] - 1.0) <= 1e-12
    e = oracle.estimate_alpha(sp.csr_matrix(np.tile(np.array([2, 3, 4], dtype=np.float32), (3, 1))))
    assert e["fell_back"] and e["n_genes_used"] == 3 and e["alpha"] == 0.25
    with pytest.raises(ValueError):
        oracle.estimate_alpha(sp.csr_matrix(np.array([[1, np.nan, 3], [2, 1, 0]], dtype=np.float32)))


def _truncated_alpha(x) -> float:
    """The reference's pre-fix estimator (a var > mean pre-filter before the
    median), planted here to show the fixture arms reject it."""
    means, var = oracle.column_moments(x)
    keep = (means > oracle.ALPHA_MU_MIN) & (var > means)
    return float(np.median((var[keep] - means[keep]) / means[keep] ** 2))


def test_alpha_simulated_fixtures():
    # pflog_reference_tests.rs:39-256 over the counts of pflog_reference_values.rs
    g = np.load(GOLDEN)
    a = sp.csr_matrix(g["A_counts"].astype(np.float32))
    e = oracle.estimate_alpha(a)
    assert e["n_genes_used"] == int(g["A_N_GENES_POOLED"]) and not e["fell_back"]
    true_a = float(g["A_ALPHA_TRUE"])
    assert abs(e["alpha"] - true_a) * float(g["A_SEPARATION"]) < abs(float(g["A_TRUNCATED_ALPHA"]) - true_a)
    assert abs(e["alpha"] - true_a) / true_a <= float(g["A_REL_BAR"])
    # the planted pre-fix estimator reproduces the fixture's truncated value and fails the separation arm
    t = _truncated_alpha(a)
    assert abs(t - float(g["A_TRUNCATED_ALPHA"])) <= 1e-12
    assert not abs(t - true_a) * float(g["A_SEPARATION"]) < abs(float(g["A_TRUNCATED_ALPHA"]) - true_a)

    b = sp.csr_matrix(g["B_counts"].astype(np.float32))
    e = oracle.estimate_alpha(b)
    assert not e["fell_back"] and e["n_genes_used"] == int(g["B_N_GENES_POOLED"])
    assert abs(e["alpha"] - float(g["B_ALPHA_TRUE"])) / float(g["B_ALPHA_TRUE"]) <= float(g["B_REL_BAR"])

    c = sp.csr_matrix(g["C_counts"].astype(np.float32))
    e = oracle.estimate_alpha(c)
    assert e["fell_back"] and e["alpha"] == 0.25 and e["n_genes_used"] == int(g["C_N_GENES_POOLED"])
    assert abs(_truncated_alpha(c) - float(g["C_TRUNCATED_ALPHA"])) <= 1e-12
