"""GPU parity on the shapes and edge cases the kernels branch on.

Each matrix goes through the C ABI and is compared with the CPU oracle on the
same input:
  * ragged: 1,000 rows (last tile partial) x 300 genes, with empty rows, a row
    whose stored values are all zero (row total 0: factor 1, the reference's
    log1p-only branch), genes never expressed, a row holding every gene (a
    forward window longer than one 64-entry batch) and a gene expressed in
    every row (a transposed-SpMM wave list longer than its prefetch capacity);
  * dense60: 600 x 200 with 60 genes expressed in every row (transposed-SpMM
    runs past the prefetched batches in every gene slot, forward windows
    longer than one batch);
  * wide3k: 3,000 genes (many gene chunks);
  * wide9k: 9,000 genes (35 forward gene chunks);
  * wide9k_general: the same matrix with SCXPCA_SPLIT_MAX_BYTES=0 (no split
    offsets: the general tile-CSC build, the tile-CSC column statistics and
    the row-parallel forward SpMM);
  * tiny: 40 x 30.
Tolerances are those of tests/test_gpu_parity.py: SpMM panels to 2e-5 of the
f64 magnitude, column statistics to 5e-7 relative, PCA singular values to 1e-4
relative and top principal angles to 1e-3 rad against the f64 oracle run with
the same Omega.
"""
from __future__ import annotations

import os
import sys

import numpy as np
import pytest

import oracle
from tests.pca_metrics import max_principal_angle, sigma_rel_err

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


class Csr:
    def __init__(self, indptr, indices, data, shape):
        self.indptr = np.ascontiguousarray(indptr, dtype=np.int64)
        self.indices = np.ascontiguousarray(indices, dtype=np.uint32)
        self.data = np.ascontiguousarray(data, dtype=np.float32)
        self.shape = shape

    def view(self):
        return oracle.CsrView(self.indptr, self.indices, self.data, self.shape)


def structured(n, nv, seed, n_types=4, density=0.15):
    """Poisson counts from n_types cell programs (a low-rank signal so the top
    components are well separated), thinned to about `density`."""
    rng = np.random.default_rng(seed)
    types = rng.integers(0, n_types, n)
    prog = rng.lognormal(-1.0, 1.2, (n_types, nv))
    lib = rng.lognormal(0.0, 0.3, n)[:, None]
    lam = prog[types] * lib
    counts = rng.poisson(lam).astype(np.float32)
    keep = rng.random((n, nv)) < density * 2.0
    counts *= keep
    return counts


def to_csr(dense):
    n, nv = dense.shape
    rows, cols = np.nonzero(dense)
    indptr = np.zeros(n + 1, np.int64)
    np.add.at(indptr, rows + 1, 1)
    indptr = np.cumsum(indptr)
    return Csr(indptr, cols, dense[rows, cols], (n, nv))


def ragged_matrix():
    d = structured(1000, 300, seed=11)
    d[:, 250:] = 0.0                 # genes never expressed
    d[5] = 0.0                       # empty rows
    d[100:111] = 0.0
    d[7] = 1.0 + (np.arange(300) % 5)  # a row holding every gene
    d[7, 250:] = 0.0
    d[:, 0] = 1.0 + (np.arange(1000) % 7)  # a gene in every row
    d[5, 0] = 0.0
    d[100:111, 0] = 0.0
    c = to_csr(d)
    # a row whose stored entries are all zero: row total 0 -> factor 1
    r = 42
    a, b = c.indptr[r], c.indptr[r + 1]
    assert b > a
    c.data[a:b] = 0.0
    return c


def dense60_matrix():
    d = structured(600, 200, seed=15, density=0.05)
    d[:, 20:80] = 1.0 + (np.arange(600)[:, None] + np.arange(60)[None, :]) % 9
    return to_csr(d)


MATRICES = {
    "ragged": ragged_matrix,
    "dense60": dense60_matrix,
    "wide3k": lambda: to_csr(structured(1500, 3000, seed=12, density=0.05)),
    "wide9k": lambda: to_csr(structured(1200, 9000, seed=13, density=0.03)),
    "wide9k_general": lambda: to_csr(structured(1200, 9000, seed=13, density=0.03)),
    "tiny": lambda: to_csr(structured(40, 30, seed=14, density=0.4)),
}


ENV = {"wide9k_general": {"SCXPCA_SPLIT_MAX_BYTES": "0"}}


@pytest.fixture(scope="module", params=list(MATRICES))
def case(request, gpu):
    name = request.param
    csr = MATRICES[name]()
    say(f"case: Matrix({name}) {csr.shape} nnz={len(csr.data)}")
    env = ENV.get(name, {})
    saved = {k: os.environ.get(k) for k in env}
    os.environ.update(env)
    try:
        mat = gpu.Matrix(csr.indptr, csr.indices, csr.data, csr.shape, target_sum=1e4, log1p=True)
    finally:
        for k, v in saved.items():
            if v is None:
                os.environ.pop(k, None)
            else:
                os.environ[k] = v
    view = csr.view()
    xt = view.with_data(oracle.transform(view, 1e4, True))
    mu = oracle.col_sums_and_sum_sq(xt)[0] / csr.shape[0]
    yield name, csr, mat, xt, mu
    mat.close()


def test_stats(case):
    name, csr, mat, xt, _ = case
    s, ss, cnt, rf = mat.stats()
    rs = oracle.row_sums(csr.view())
    np.testing.assert_allclose(rf, np.where(rs > 0, 1e4 / np.where(rs > 0, rs, 1.0), 1.0), rtol=1e-15)
    os_, oss = oracle.col_sums_and_sum_sq(xt)
    np.testing.assert_allclose(s, os_, rtol=5e-7, atol=1e-9)
    np.testing.assert_allclose(ss, oss, rtol=5e-7, atol=1e-9)
    np.testing.assert_array_equal(cnt, np.bincount(csr.indices, minlength=csr.shape[1]))


@pytest.mark.parametrize("k", [60, 100, 5])
def test_spmm_forward(case, k):
    name, csr, mat, xt, mu = case
    say(f"case: {name} forward k={k}")
    m = np.random.default_rng(k).standard_normal((csr.shape[1], k))
    ref = oracle.spmm_forward(xt, m, mu)
    got = mat.spmm(m, transpose=False, means=mu)
    assert np.abs(got - ref).max() / np.abs(ref).max() < 2e-5


@pytest.mark.parametrize("k", [60, 100, 5])
def test_spmm_transpose(case, k):
    name, csr, mat, xt, mu = case
    say(f"case: {name} transpose k={k}")
    q = np.random.default_rng(100 + k).standard_normal((csr.shape[0], k))
    ref = oracle.spmm_transpose(xt, q, mu)
    got = mat.spmm(q, transpose=True, means=mu)
    assert np.abs(got - ref).max() / np.abs(ref).max() < 2e-5


def _bf16(x):
    """Round to bf16 (round to nearest even), as f64."""
    u = np.ascontiguousarray(x, dtype=np.float32).view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32).astype(np.float64)


@pytest.mark.parametrize("k", [60, 100, 5])
def test_spmm_bf16(case, k):
    """The BF16 sketch form's products: x~ rounded to bf16 (the packed
    entries), times the f32 panel with f32 accumulation. Both directions
    against the f64 product of bf16-rounded X~ and the panel, 2e-5 of the f64
    magnitude (f32 accumulation and the centering's cancellation, the FP32
    form's stage bar), and against the FP32 form's product to the bf16
    rounding of X~ (2^-9 per value; 1e-2 of the magnitude after the
    centering's cancellation)."""
    name, csr, mat, xt, mu = case
    xb = xt.with_data(_bf16(xt.data).astype(np.float32))
    rng = np.random.default_rng(200 + k)
    m = rng.standard_normal((csr.shape[1], k))
    q = rng.standard_normal((csr.shape[0], k))
    say(f"case: {name} bf16 forward + transpose k={k}")
    got_f = mat.spmm(m, transpose=False, means=mu, bf16=True)
    got_t = mat.spmm(q, transpose=True, means=mu, bf16=True)
    m32 = m.astype(np.float32).astype(np.float64)
    q32 = q.astype(np.float32).astype(np.float64)
    ref_f = oracle.spmm_forward(xb, m32, mu)
    ref_t = oracle.spmm_transpose(xb, q32, mu)
    ef = np.abs(got_f - ref_f).max() / np.abs(ref_f).max()
    et = np.abs(got_t - ref_t).max() / np.abs(ref_t).max()
    f32_f = oracle.spmm_forward(xt, m, mu)
    f32_t = oracle.spmm_transpose(xt, q, mu)
    say(f"{name} bf16 k={k}: fwd {ef:.2e} t {et:.2e}; vs fp32 form "
        f"{np.abs(got_f - f32_f).max() / np.abs(f32_f).max():.2e} / {np.abs(got_t - f32_t).max() / np.abs(f32_t).max():.2e}")
    assert ef < 2e-5 and et < 2e-5
    assert np.abs(got_f - f32_f).max() <= 1e-2 * np.abs(f32_f).max()
    assert np.abs(got_t - f32_t).max() <= 1e-2 * np.abs(f32_t).max()


def test_pca(case):
    name, csr, mat, xt, mu = case
    n, nv = csr.shape
    n_comps, p = (5, 3) if name == "tiny" else (20, 10)
    say(f"case: {name} pca n_comps={n_comps}")
    ref = oracle.randomized_pca(csr.view(), n_comps=n_comps, n_oversamples=p, n_power_iterations=3, seed=0,
                                target_sum=1e4, log1p=True)
    got = mat.pca(n_comps=n_comps, n_oversamples=p, n_power_iterations=3, seed=0, qr_method=1)
    top = 3
    se = sigma_rel_err(ref["sigma_all"][:top], got["sigma"][:top])
    ang = max_principal_angle(ref["components"].T[:, :top], got["components"].T[:, :top].astype(np.float64))
    say(f"{name}: sigma rel err top{top} {se.max():.2e}, angle {ang:.2e}")
    assert se.max() <= 1e-4
    assert ang <= 1e-3
    np.testing.assert_allclose(got["mean"], ref["mean"], rtol=1e-6, atol=1e-12)
