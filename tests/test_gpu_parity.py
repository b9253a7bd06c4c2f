"""GPU parity: the HIP path through the C ABI against the CPU oracle.

Tolerances: FP32 device arithmetic against the f64 oracle. Stage outputs are
compared relative to the magnitude of the f64 result (2e-5 for SpMM panels, a
few f32 ulps for the X~ column statistics); full PCA results to the contract
of the path: singular values <= 1e-4 relative, largest principal angle <= 1e-3
rad, per-component scores <= 1e-3 relative after sign alignment.
"""
from __future__ import annotations

import sys

import numpy as np
import pytest

import oracle
from scx_amd import synth
from tests.pca_metrics import max_principal_angle, score_rel_err, sigma_rel_err

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def small():
    # the first 3000 rows of the 10k x 2k workload: real NB counts, ~10% density
    return synth.generate(synth.CFG1, n_obs=3000, workers=1)


@pytest.fixture(scope="module")
def small_mat(gpu, small):
    say("case: Matrix(cfg1 head 3000)")
    return gpu.Matrix(small.indptr, small.indices, small.data, small.shape, target_sum=1e4, log1p=True)


@pytest.mark.parametrize("rows,cols,seed", [(2000, 60, 0), (300, 7, 42), (33538, 120, 0), (5000, 128, 12345)])
def test_device_omega_matches_oracle_bit_exactly(gpu, rows, cols, seed):
    say(f"case: device Omega {rows}x{cols} seed={seed}")
    got = gpu.gaussian_device(rows, cols, seed)
    ref = oracle.random_gaussian(rows, cols, seed)
    np.testing.assert_array_equal(got.view(np.uint64), ref.view(np.uint64))


def test_stats_parity(small, small_mat):
    s, ss, cnt, rf = small_mat.stats()
    view = oracle.CsrView(small.indptr, small.indices, small.data, small.shape)
    rs = oracle.row_sums(view)
    np.testing.assert_allclose(rf, np.where(rs > 0, 1e4 / rs, 1.0), rtol=1e-15)
    xt = oracle.transform(view, 1e4, True)
    os_, oss = oracle.col_sums_and_sum_sq(view.with_data(xt))
    np.testing.assert_allclose(s, os_, rtol=5e-7, atol=1e-9)
    np.testing.assert_allclose(ss, oss, rtol=5e-7, atol=1e-9)
    np.testing.assert_array_equal(cnt, np.bincount(small.indices, minlength=small.shape[1]))


@pytest.mark.parametrize("k,centered", [(60, True), (60, False), (100, True), (7, True)])
def test_spmm_forward_parity(small, small_mat, k, centered):
    say(f"case: spmm forward k={k} centered={centered}")
    rng = np.random.default_rng(k)
    m = rng.standard_normal((small.shape[1], k))
    view = oracle.CsrView(small.indptr, small.indices, small.data, small.shape)
    xt = view.with_data(oracle.transform(view, 1e4, True))
    mu = oracle.col_sums_and_sum_sq(xt)[0] / small.shape[0] if centered else None
    ref = oracle.spmm_forward(xt, m, mu)
    got = small_mat.spmm(m, transpose=False, means=mu)
    scale = np.abs(ref).max()
    assert np.abs(got - ref).max() / scale < 2e-5


@pytest.mark.parametrize("k,centered", [(60, True), (60, False), (120, True), (3, True)])
def test_spmm_transpose_parity(small, small_mat, k, centered):
    say(f"case: spmm transpose k={k} centered={centered}")
    rng = np.random.default_rng(100 + k)
    q = rng.standard_normal((small.shape[0], k)) / np.sqrt(small.shape[0])
    view = oracle.CsrView(small.indptr, small.indices, small.data, small.shape)
    xt = view.with_data(oracle.transform(view, 1e4, True))
    mu = oracle.col_sums_and_sum_sq(xt)[0] / small.shape[0] if centered else None
    ref = oracle.spmm_transpose(xt, q, mu)
    got = small_mat.spmm(q, transpose=True, means=mu)
    scale = np.abs(ref).max()
    assert np.abs(got - ref).max() / scale < 2e-5


@pytest.mark.parametrize("method", [0, 1])
@pytest.mark.parametrize("rows,k", [(5000, 60), (20000, 120), (300, 5)])
def test_qr(gpu, method, rows, k):
    say(f"case: qr rows={rows} k={k} method={method}")
    rng = np.random.default_rng(rows + k)
    a = rng.standard_normal((rows, k)) @ np.diag(np.logspace(0, 2, k))
    q = gpu.qr(a, qr_method=method)
    assert np.abs(q.T @ q - np.eye(k)).max() < 1e-4
    assert max_principal_angle(a, q) < 1e-4


def _oracle_pca(csr, w, n_comps, n_over, n_iter, seed=0):
    return oracle.randomized_pca(oracle.CsrView(csr.indptr, csr.indices, csr.data, csr.shape), n_comps=n_comps,
                                 n_oversamples=n_over, n_power_iterations=n_iter, zero_center=True, seed=seed,
                                 target_sum=w.target_sum, log1p=True)


@pytest.mark.parametrize("n_iter,qr_method", [(2, 0), (3, 1), (3, 0), (0, 0), (0, 1), (1, 1)])
def test_pca_parity_small(small, small_mat, n_iter, qr_method):
    say(f"case: pca small n_iter={n_iter} qr={qr_method}")
    ref = _oracle_pca(small, synth.CFG1, 50, 10, n_iter)
    got = small_mat.pca(n_comps=50, n_oversamples=10, n_power_iterations=n_iter, seed=0, qr_method=qr_method)
    se = sigma_rel_err(ref["sigma_all"][:50], got["sigma"][:50])
    say("sigma rel err max", se.max())
    assert se.max() <= 1e-4
    ang = max_principal_angle(ref["components"].T[:, :10], got["components"].T[:, :10].astype(np.float64))
    say("angle top10", ang)
    assert ang <= 1e-3
    np.testing.assert_allclose(got["mean"], ref["mean"], rtol=1e-6, atol=1e-12)
    np.testing.assert_allclose(got["variance_ratio"], ref["variance_ratio"], rtol=2e-4)
    assert abs(got["total_variance"] - (ref["variance"] / ref["variance_ratio"])[0]) <= 1e-6 * got["total_variance"]


@pytest.mark.parametrize("on_device", [False, True])
def test_repeated_column_in_a_row_is_rejected(gpu, on_device):
    # the transposed SpMM's tile layout holds one entry per (row, gene): a row
    # that repeats a column index is refused with a clear error, on both paths
    indptr = np.array([0, 2, 5, 6], dtype=np.int64)
    indices = np.array([0, 3, 1, 1, 4, 2], dtype=np.uint32)
    data = np.ones(6, dtype=np.float32)
    say(f"case: repeated column, on_device={on_device}")
    if on_device:
        import torch

        dev = torch.device("cuda:0")
        t = (torch.from_numpy(indptr).to(dev), torch.from_numpy(indices.view(np.int32)).to(dev),
             torch.from_numpy(data).to(dev))
        torch.cuda.synchronize()
        with pytest.raises(ValueError, match="repeats a column index"):
            gpu.Matrix(*t, (3, 5), target_sum=1e4, log1p=True, device=0)
    else:
        with pytest.raises(ValueError, match="repeats a column index"):
            gpu.Matrix(indptr, indices, data, (3, 5), target_sum=1e4, log1p=True)


def _device_csr(rows, n_vars):
    """Device tensors of a CSR whose rows hold the given column lists."""
    import torch

    indptr = np.zeros(len(rows) + 1, dtype=np.int64)
    indptr[1:] = np.cumsum([len(r) for r in rows])
    indices = np.concatenate([np.asarray(r, dtype=np.uint32) for r in rows])
    data = np.ones(len(indices), dtype=np.float32)
    dev = torch.device("cuda:0")
    t = (torch.from_numpy(indptr).to(dev), torch.from_numpy(indices.view(np.int32)).to(dev),
         torch.from_numpy(data).to(dev))
    torch.cuda.synchronize()
    return t, (len(rows), n_vars)


@pytest.mark.parametrize("pos", [10, 64, 128, 256])
@pytest.mark.parametrize("kind", ["decrease", "repeat"])
def test_row_order_checked_across_batches(gpu, kind, pos):
    # the device check compares each entry with its predecessor: inside a
    # 64-entry batch (pos 10) and across batches, where the predecessor is the
    # previous batch's last entry (pos 64, 128; pos 256 in a row of 300
    # entries, the long-row path); every other row is valid
    cols = list(range(0, 2 * 300, 2))  # 300 sorted entries, n_vars 600
    bad = list(cols)
    bad[pos] = bad[pos - 1] - (1 if kind == "decrease" else 0)
    rows = [cols[:130], bad[: max(pos + 2, 130)], cols[:5]]
    t, shape = _device_csr(rows, 600)
    say(f"case: row order {kind} at {pos}")
    msg = "sorted column indices" if kind == "decrease" else "repeats a column index"
    with pytest.raises(ValueError, match=msg):
        gpu.Matrix(*t, shape, target_sum=1e4, log1p=True, device=0)


def test_valid_long_rows_accepted(gpu):
    # rows of 63, 64, 65, 130, 256, 257 and 300 sorted entries (batch and
    # long-row boundaries) pass the device check, and their column counts come back
    rows = [list(range(0, 2 * n, 2)) for n in (63, 64, 65, 130, 256, 257, 300)]
    t, shape = _device_csr(rows, 600)
    say("case: valid long rows")
    m = gpu.Matrix(*t, shape, target_sum=1e4, log1p=True, device=0)
    try:
        _, _, cnt, _ = m.stats()
        want = np.bincount(np.concatenate([np.asarray(r) for r in rows]), minlength=600)
        np.testing.assert_array_equal(cnt, want)
    finally:
        m.close()


@pytest.mark.parametrize("value", ["lots", "-1", "1e9x"])
def test_invalid_split_cap_is_rejected(gpu, small, monkeypatch, value):
    # SCXPCA_SPLIT_MAX_BYTES is read when a matrix is created; a value that is
    # not a non-negative number is refused, not silently ignored
    say(f"case: SCXPCA_SPLIT_MAX_BYTES={value!r}")
    monkeypatch.setenv("SCXPCA_SPLIT_MAX_BYTES", value)
    with pytest.raises(ValueError, match="SCXPCA_SPLIT_MAX_BYTES"):
        gpu.Matrix(small.indptr, small.indices, small.data, small.shape, target_sum=1e4, log1p=True)


def test_rank_deficient_matrix(gpu):
    """k above the rank (10 distinct cells, each repeated 50 times): the stable
    mode returns the oracle's (Householder) spectrum, zeros beyond the rank;
    CholeskyQR2 reports its non-SPD error, as the reference's cuSOLVER route
    does (scx-gpu/src/cusolver.rs:381-385)."""
    import scipy.sparse as sp

    rng = np.random.default_rng(5)
    base = rng.negative_binomial(2, 0.4, size=(10, 40)).astype(np.float32)
    x = sp.csr_matrix(np.repeat(base, 50, axis=0))
    say("case: rank-deficient 500x40 rank<=10, k=30")
    ref = oracle.randomized_pca(oracle.CsrView(x.indptr, x.indices, x.data, x.shape), n_comps=20, n_oversamples=10,
                                n_power_iterations=2, zero_center=True, seed=0, target_sum=1e4, log1p=True)
    m = gpu.Matrix(x.indptr.astype(np.int64), x.indices.astype(np.uint32), x.data, x.shape, target_sum=1e4,
                   log1p=True)
    try:
        got = m.pca(n_comps=20, n_oversamples=10, n_power_iterations=2, seed=0, qr_method=0)
        with pytest.raises(gpu.LinAlgError, match="CholeskyQR2 failed: non-SPD"):
            m.pca(n_comps=20, n_oversamples=10, n_power_iterations=2, seed=0, qr_method=1)
    finally:
        m.close()
    s_ref, s_got = ref["sigma_all"], got["sigma"]
    say("rank-deficient sigma", s_got[:11], "ref", s_ref[:11])
    assert np.all(sigma_rel_err(s_ref[:9], s_got[:9]) <= 1e-4)
    assert np.all(s_got[9:] <= 1e-4 * s_got[0]) and np.all(s_ref[9:] <= 1e-9 * s_ref[0])
    assert max_principal_angle(ref["components"].T[:, :9], got["components"].T[:, :9].astype(np.float64)) <= 1e-3
    np.testing.assert_allclose(got["variance_ratio"][:9], ref["variance_ratio"][:9], rtol=2e-4)
    assert np.all(np.isfinite(got["embeddings"])) and np.all(np.isfinite(got["components"]))


def test_repeated_calls_are_deterministic(small, small_mat):
    # the same call on the same resident matrix gives the same bits, also with
    # calls of another seed, k, precision and a refresh (re-prepare and gene
    # re-deal) in between; the other seed's result differs
    say("case: repeated pca calls")
    kw = dict(n_comps=30, n_oversamples=10, n_power_iterations=3, seed=7, qr_method=1)
    a = small_mat.pca(**kw)
    other = small_mat.pca(**dict(kw, seed=8))
    small_mat.pca(**dict(kw, n_comps=20, n_oversamples=5))
    small_mat.pca(**dict(kw, precision=1))
    small_mat.refresh()
    b = small_mat.pca(**kw)
    for key in ("sigma", "embeddings", "components", "variance_ratio"):
        np.testing.assert_array_equal(a[key], b[key])
    assert not np.array_equal(a["sigma"], other["sigma"])
