"""CPU tests that pin the oracle (no GPU).

Each check is chosen so that a plausible mistake in the restatement (wrong
round count, word order, index, sign, dropped term) fails it:
  * reference KAT: seeded_permutation(16, 42) (scx-ops/src/shuffle_order_tests.rs:67-80)
  * published ChaCha8 keystream for the all-zero 256-bit key and nonce
  * published leading entries of rand_distr's ZIG_NORM_X / ZIG_NORM_F and the
    ziggurat's defining invariants (equal box areas, F = pdf(X))
  * distribution of the StandardNormal stream (moments, KS, tail mass)
  * the reference fixtures test_csr_10x5 (cpu.rs:2326) and cancellation_csr
    (cpu.rs:2201) with their analytic answers
  * brute-force dense f64 products and LAPACK SVD/QR for the SpMM / QR / SVD /
    PCA stages.
"""
from __future__ import annotations

import numpy as np
import pytest
import scipy.stats as st

import oracle
from scx_amd import synth
from tests.pca_metrics import dense_pca, dense_xtilde, max_principal_angle, score_rel_err


# ---------------------------------------------------------------- RNG
def test_shuffle_kat_pins_chacha8_seeding_and_gen_range():
    assert oracle.seeded_permutation(16, 42).tolist() == [0, 4, 2, 12, 3, 13, 11, 5, 15, 8, 6, 7, 14, 9, 10, 1]


def test_chacha8_zero_key_keystream():
    # ChaCha, 8 rounds, 256-bit all-zero key, all-zero 64-bit IV, block 0
    # (the standard published ChaCha8 test vector).
    out = oracle.chacha8_block(np.zeros(8, np.uint32)).astype("<u4").tobytes().hex()
    assert out == ("3e00ef2f895f40d67f5bb8e81f09a5a12c840ec3ce9a7f3b181be188ef711a1e"
                   "984ce172b9216f419f445367456d5619314a42a3da86b001387bfdb80e0cfe42")


def test_chacha8_counter_and_refill_order():
    # the u32 stream is four consecutive blocks per refill, block after block
    seed = 7
    s = oracle.chacha8_u32(seed, 256)
    r = oracle.chacha8_u32(seed, 1)  # noqa: F841 - same seed is deterministic
    # key: PCG32 expansion of the seed
    key = _pcg32_key(seed)
    for b in range(16):
        np.testing.assert_array_equal(s[16 * b:16 * b + 16], oracle.chacha8_block(key, counter=b))


def _pcg32_key(seed: int) -> np.ndarray:
    # rand_core 0.6.4 seed_from_u64, written independently in Python integers
    mask = (1 << 64) - 1
    st_ = seed
    out = []
    for _ in range(8):
        st_ = (st_ * 6364136223846793005 + 11634580027462260723) & mask
        xs = (((st_ >> 18) ^ st_) >> 27) & 0xFFFFFFFF
        rot = st_ >> 59
        out.append(((xs >> rot) | (xs << ((32 - rot) & 31))) & 0xFFFFFFFF)
    return np.array(out, dtype=np.uint32)


def test_next_u64_is_low_word_first_pairs():
    # Ω only ever reads u64s, so reads never straddle the buffer end: the
    # u64 stream is the u32 stream paired low word first.
    u32 = oracle.chacha8_u32(3, 400).astype(np.uint64)
    u64 = oracle.chacha8_u64(3, 200)
    np.testing.assert_array_equal(u64, u32[0::2] | (u32[1::2] << np.uint64(32)))


def test_ziggurat_tables():
    x, f = oracle.zig_norm_tables()
    # leading entries as printed in rand_distr's ziggurat_tables.rs
    assert "%.18f" % x[0] == "3.910757959537090045"
    assert "%.18f" % x[1] == "3.654152885361008796"
    assert "%.18f" % x[2] == "3.449278298560964462"
    assert "%.18f" % x[3] == "3.320244733839166074"
    assert "%.18f" % f[0] == "0.000477467764586655"
    assert "%.18f" % f[1] == "0.001260285930498598"
    assert "%.18f" % f[2] == "0.002609072746106363"
    assert "%.18f" % f[3] == "0.004037972593371872"
    assert x[256] == 0.0 and f[256] == 1.0
    assert np.all(np.diff(x) < 0) and np.all(np.diff(f) > 0)
    # F_i = pdf of the unrounded x_i, then rounded: within an ulp or two
    np.testing.assert_allclose(f, np.exp(-x * x / 2), rtol=5e-16, atol=1e-19)
    # equal box areas V = x_i (f_{i+1} - f_i): exact by the recurrence for the
    # boxes it builds; the top cap closes at f(0) = 1 only to the precision of
    # the 12-digit V (Doornik's constants)
    area = x[1:256] * (f[2:257] - f[1:256])
    np.testing.assert_allclose(area[:-1], 0.00492867323399, rtol=2e-11)
    np.testing.assert_allclose(area[-1], 0.00492867323399, rtol=1e-8)
    # base strip: x0 = V / f(R)
    assert abs(x[0] * f[1] - 0.00492867323399) < 1e-15


def test_standard_normal_distribution():
    g = oracle.random_gaussian(2000, 1000, 11).ravel()
    assert abs(g.mean()) < 5e-3
    assert abs(g.var() - 1.0) < 5e-3
    assert st.kstest(g, "norm").pvalue > 1e-3
    # tail beyond R (the zero_case branch): P(|x| > R) = 2 * Phi(-R)
    tail = np.mean(np.abs(g) > 3.6541528853610088)
    expect = 2 * st.norm.sf(3.6541528853610088)
    assert abs(tail - expect) < 6 * np.sqrt(expect / g.size)
    assert st.skew(g) < 0.01 and abs(st.kurtosis(g)) < 0.02


def test_random_gaussian_is_one_stream_row_major():
    a = oracle.random_gaussian(30, 7, 5)
    b = oracle.random_gaussian(1, 210, 5)
    np.testing.assert_array_equal(a.ravel(), b.ravel())


def test_product_omega_matches_oracle_bit_exactly():
    from scx_amd import _lib

    for seed in (0, 1, 42, 2**64 - 1):
        got = _lib.gaussian(257, 61, seed)
        ref = oracle.random_gaussian(257, 61, seed)
        assert np.array_equal(got.view(np.uint64), ref.view(np.uint64)), seed


# ---------------------------------------------------------------- fixtures
def csr_10x5():
    """test_csr_10x5 (scx-accel/src/pca/cpu.rs:2326)."""
    indptr = np.array([0, 2, 4, 7, 9, 12, 14, 17, 19, 22, 24])
    indices = np.array([0, 2, 1, 3, 0, 2, 4, 1, 4, 0, 2, 3, 1, 3, 0, 2, 4, 1, 3, 0, 2, 4, 1, 3])
    data = np.array([1.0, 3.0, 2.0, 4.0, 5.0, 1.0, 2.0, 3.0, 6.0, 2.0, 4.0, 1.0, 1.0, 3.0, 3.0, 2.0,
                     5.0, 4.0, 2.0, 1.0, 5.0, 3.0, 2.0, 1.0])
    return synth.Csr(indptr.astype(np.int64), indices.astype(np.uint32), data.astype(np.float32), (10, 5))


def cancellation_csr(n_rows: int):
    """cancellation_csr (cpu.rs:2201): col 0 = 1e6 +/- 1, col 1 = r % 5, 3 columns."""
    idx, val, indptr = [], [], [0]
    for r in range(n_rows):
        idx += [0, 1]
        val += [np.float32(1.0e6) + (1.0 if r % 2 == 0 else -1.0), float(r % 5)]
        indptr.append(len(idx))
    return synth.Csr(np.array(indptr, np.int64), np.array(idx, np.uint32), np.array(val, np.float32), (n_rows, 3))


def _view(c):
    return oracle.CsrView(c.indptr, c.indices, c.data, c.shape)


def _close_ulp(got_f32, ref_f64, ulps=1):
    ref32 = ref_f64.astype(np.float32)
    assert np.all(np.abs(got_f32 - ref32) <= ulps * np.spacing(np.abs(ref32))), (got_f32, ref32)


# ---------------------------------------------------------------- X~
def test_transform_matches_reference_formula():
    c = csr_10x5()
    got = oracle.transform(_view(c), 1e4, True)
    rs = np.add.reduceat(c.data.astype(np.float64), c.indptr[:-1])
    rows = np.repeat(np.arange(10), np.diff(c.indptr))
    scaled = (c.data.astype(np.float64) * (1e4 / rs)[rows]).astype(np.float32)
    # the oracle calls glibc log1pf (what Rust's f32::ln_1p resolves to on
    # Linux); compare with the f64 log1p of the same f32 input to 1 f32 ulp
    _close_ulp(got, np.log1p(scaled.astype(np.float64)))
    # log1p only (fused_ops.rs:658-676 values ln(6), ln(11))
    lone = oracle.transform(oracle.CsrView([0, 2], [0, 1], [5.0, 10.0], (1, 2)), None, True)
    assert abs(lone[0] - np.log(6.0)) < 1e-6 and abs(lone[1] - np.log(11.0)) < 1e-6


def test_transform_nonpositive_row_total_applies_log1p_only():
    # transforms.rs:40-72: rows summing to <= 0 skip normalization but keep log1p
    v = oracle.CsrView([0, 2, 4, 4], [0, 1, 0, 1], [0.5, -0.5, -0.25, 0.2], (3, 2))
    got = oracle.transform(v, 1e4, True)
    _close_ulp(got, np.log1p(np.array([0.5, -0.5, -0.25, 0.2], np.float64)))
    got2 = oracle.transform(v, 1e4, False)
    np.testing.assert_array_equal(got2, np.array([0.5, -0.5, -0.25, 0.2], np.float32))


def test_col_stats_and_row_sums_closed_form():
    c = csr_10x5()
    d = c.to_scipy().toarray().astype(np.float64)
    s, ss = oracle.col_sums_and_sum_sq(_view(c))
    np.testing.assert_array_equal(s, d.sum(0))
    np.testing.assert_array_equal(ss, (d * d).sum(0))
    np.testing.assert_array_equal(oracle.row_sums(_view(c)), d.sum(1))


# ---------------------------------------------------------------- SpMM
@pytest.mark.parametrize("centered", [True, False])
def test_spmm_matches_dense(centered):
    # test_spmm_matches_dense (cpu.rs:2415), widened to a random 300 x 120 CSR
    c = synth.random_csr(300, 120, 0.15, seed=4, integer=False)
    d = c.to_scipy().toarray().astype(np.float64)
    mu = d.mean(0) if centered else None
    m = np.random.default_rng(1).standard_normal((120, 17))
    y = oracle.spmm_forward(_view(c), m, mu)
    ref = (d - (mu if centered else 0)) @ m
    np.testing.assert_allclose(y, ref, rtol=0, atol=1e-10 * np.abs(ref).max())
    q = np.random.default_rng(2).standard_normal((300, 17))
    z = oracle.spmm_transpose(_view(c), q, mu)
    ref_t = (d - (mu if centered else 0)).T @ q
    np.testing.assert_allclose(z, ref_t, rtol=0, atol=1e-10 * np.abs(ref_t).max())


def test_spmm_transpose_bits_independent_of_threads():
    # output partitioning: every element sums rows in ascending order
    c = synth.random_csr(2000, 333, 0.1, seed=9, integer=False)
    q = np.random.default_rng(3).standard_normal((2000, 23))
    mu = np.random.default_rng(4).random(333)
    z1 = oracle.spmm_transpose(_view(c), q, mu, nthreads=1)
    for t in (2, 3, 8):
        assert np.array_equal(oracle.spmm_transpose(_view(c), q, mu, nthreads=t), z1)
    y1 = oracle.spmm_forward(_view(c), q[:333], mu, nthreads=1)
    assert np.array_equal(oracle.spmm_forward(_view(c), q[:333], mu, nthreads=8), y1)


def test_spmm_unsorted_rows_same_values():
    c = synth.random_csr(200, 50, 0.2, seed=5, integer=False)
    u = synth.random_csr(200, 50, 0.2, seed=5, integer=False, sorted_rows=False)
    q = np.random.default_rng(6).standard_normal((200, 5))
    np.testing.assert_allclose(oracle.spmm_transpose(_view(u), q), oracle.spmm_transpose(_view(c), q), rtol=1e-12)


# ---------------------------------------------------------------- QR / SVD
def test_qr_orthonormal_and_span():
    a = np.random.default_rng(0).standard_normal((500, 40)) @ np.diag(np.logspace(0, 3, 40))
    q = oracle.qr_thin_q(a)
    assert np.abs(q.T @ q - np.eye(40)).max() < 1e-13
    assert max_principal_angle(a, q) < 1e-10
    # the test_qr_orthonormal fixture (cpu.rs:2447)
    q2 = oracle.qr_thin_q(np.array([[1.0, 0.0], [0.0, 1.0], [1.0, 1.0], [0.0, 0.0]]))
    assert np.abs(q2.T @ q2 - np.eye(2)).max() < 1e-14


def test_thin_svd_matches_lapack():
    b = np.random.default_rng(1).standard_normal((300, 25)) @ np.diag(np.logspace(0, -4, 25))
    u, s, vt = oracle.thin_svd(b)
    s_ref = np.linalg.svd(b, compute_uv=False)
    np.testing.assert_allclose(s, s_ref, rtol=1e-12)
    np.testing.assert_allclose(u @ np.diag(s) @ vt, b, atol=1e-13)
    assert np.abs(u.T @ u - np.eye(25)).max() < 1e-12
    assert np.all(np.diff(s) <= 0)


# ---------------------------------------------------------------- total variance
def test_total_variance_closed_form_and_guard():
    c = csr_10x5()
    s, ss = oracle.col_sums_and_sum_sq(_view(c))
    mu = s / 10
    tv, rec = oracle.total_variance(_view(c), ss, mu)
    d = c.to_scipy().toarray().astype(np.float64)
    assert not rec
    assert abs(tv - d.var(0, ddof=1).sum()) < 1e-12
    # cancellation fixture: analytic centered variance 3n/(n-1)
    n = 1000
    cc = cancellation_csr(n)
    s, ss = oracle.col_sums_and_sum_sq(_view(cc))
    tv, rec = oracle.total_variance(_view(cc), ss, s / n)
    assert rec
    assert abs(tv - 3 * n / (n - 1)) < 1e-9
    # uncentered: plain sum of squares / (n - 1)
    tvu, _ = oracle.total_variance(_view(c), oracle.col_sums_and_sum_sq(_view(c))[1], None)
    assert abs(tvu - (d * d).sum() / 9) < 1e-12


# ---------------------------------------------------------------- full PCA
def test_pca_errors_match_validate_inputs():
    c = csr_10x5()
    with pytest.raises(ValueError, match="exceeds matrix rank bound min"):
        oracle.randomized_pca(_view(c), n_comps=6)
    with pytest.raises(ValueError, match="n_components must be > 0"):
        oracle.randomized_pca(_view(c), n_comps=0)


def test_pca_basic_fixture_invariants():
    # test_pca_basic (cpu.rs:2342): 3 components, p=5, q=2, seed 42
    r = oracle.randomized_pca(_view(csr_10x5()), n_comps=3, n_oversamples=5, n_power_iterations=2, seed=42)
    assert r["embeddings"].shape == (10, 3) and r["components"].shape == (3, 5)
    comp = r["components"]
    assert np.abs(comp @ comp.T - np.eye(3)).max() < 1e-12
    assert np.all(np.diff(r["variance"]) <= 1e-12)
    assert r["variance_ratio"].sum() <= 1 + 1e-12
    # k = min(8, 5, 10) = 5 = full rank: the randomized SVD is exact
    d = csr_10x5().to_scipy().toarray().astype(np.float64)
    ex = dense_pca(d, 3)
    np.testing.assert_allclose(r["sigma_all"][:3], ex["sigma"], rtol=1e-12)
    assert np.abs(score_rel_err(ex["embeddings"], r["embeddings"])).max() < 1e-10


def _structured(n=1500, nv=300, seed=0):
    """Counts with 4 cell types (distinct gene programs): a decaying spectrum
    where the top components are well separated."""
    rng = np.random.default_rng(seed)
    base = rng.lognormal(-1, 1.2, nv)
    prog = np.ones((4, nv))
    for t in range(4):
        prog[t, rng.choice(nv, 40, replace=False)] *= rng.uniform(4, 12)
    cell_type = rng.integers(0, 4, n)
    lib = rng.lognormal(7, 0.4, n)
    mu = lib[:, None] * (base[None, :] * prog[cell_type]) / base.sum() * 0.5
    lam = rng.gamma(2.0, mu / 2.0)
    x = rng.poisson(lam).astype(np.float32)
    import scipy.sparse as sp

    m = sp.csr_matrix(x)
    m.sort_indices()
    return synth.Csr(m.indptr.astype(np.int64), m.indices.astype(np.uint32), m.data.astype(np.float32), m.shape)


@pytest.mark.parametrize("n_iter", [5, 8])
def test_pca_matches_dense_fp64_oracle(n_iter):
    # enough power iterations for the randomized SVD itself to converge on the
    # three well-separated cell-type components
    c = _structured()
    xd = dense_xtilde(c, 1e4, True)
    ex = dense_pca(xd, 10)
    r = oracle.randomized_pca(_view(c), n_comps=10, n_oversamples=10, n_power_iterations=n_iter, seed=0,
                              target_sum=1e4, log1p=True)
    # top-3 are well separated (cell-type programs): tight agreement
    np.testing.assert_allclose(r["sigma_all"][:3], ex["sigma"][:3], rtol=1e-6)
    assert max_principal_angle(ex["components"][:3].T, r["components"][:3].T) < 1e-4
    assert score_rel_err(ex["embeddings"][:, :3], r["embeddings"][:, :3]).max() < 1e-4
    # dense_xtilde uses numpy float32 log1p, the oracle glibc log1pf: 1-ulp apart
    np.testing.assert_allclose(r["mean"], xd.mean(0), rtol=1e-6)
    np.testing.assert_allclose(r["variance"] / r["variance_ratio"], ex["total_variance"], rtol=1e-6)


def test_pca_threads_do_not_change_bits_of_spmm_stages():
    c = _structured(600, 120)
    a = oracle.randomized_pca(_view(c), n_comps=5, seed=3, target_sum=1e4, log1p=True, nthreads=1)
    b = oracle.randomized_pca(_view(c), n_comps=5, seed=3, target_sum=1e4, log1p=True, nthreads=4)
    # SpMMs are bit-identical; the QR is column-parallel, also order-fixed
    assert np.array_equal(a["sigma_all"], b["sigma_all"])


def test_covariance_oracle_matches_dense_svd():
    """The covariance route's oracle against the SVD of the explicitly centred
    X~: eigenvalues = sigma^2 / (n - 1), embeddings = U Sigma (up to sign),
    components = right singular vectors (up to sign); and uncentred: the
    eigenvalues of X~^T X~ / (n - 1)."""
    import scipy.sparse as sp

    from scx_amd import synth
    x = synth.generate(synth.CFG1, n_obs=600, workers=1)
    m = x.to_scipy()[:, :90].tocsr()
    m.sort_indices()
    v = oracle.CsrView(m.indptr.astype(np.int64), m.indices.astype(np.uint32), m.data.astype(np.float32), m.shape)
    xt = v.with_data(oracle.transform(v, 1e4, True))
    d = sp.csr_matrix((xt.data.astype(np.float64), xt.indices, xt.indptr), shape=m.shape).toarray()
    n = m.shape[0]
    r = oracle.covariance_pca(v, n_comps=8, zero_center=True, target_sum=1e4, log1p=True)
    u, s, vt = np.linalg.svd(d - d.mean(axis=0), full_matrices=False)
    np.testing.assert_allclose(r["variance"], s[:8] ** 2 / (n - 1), rtol=1e-12)
    for j in range(8):
        sgn = np.sign(np.dot(r["components"][j], vt[j]))
        np.testing.assert_allclose(r["components"][j], sgn * vt[j], atol=1e-9)
        np.testing.assert_allclose(r["embeddings"][:, j], sgn * u[:, j] * s[j], atol=1e-9 * s[0])
    np.testing.assert_allclose(r["mean"], d.mean(axis=0), rtol=1e-13, atol=1e-15)
    assert r["variance_ratio"].sum() <= 1.0 + 1e-12
    ru = oracle.covariance_pca(v, n_comps=5, zero_center=False, target_sum=1e4, log1p=True)
    w = np.linalg.eigvalsh(d.T @ d / (n - 1))[::-1]
    np.testing.assert_allclose(ru["variance"], w[:5], rtol=1e-12)


def test_method_routing():
    """GPU: every method but an explicit "covariance" is randomized
    (pyscx/src/accel/pca.rs:384-389); the CPU router's "auto" is covariance at
    <= 5,000 genes (cpu.rs:1643-1660)."""
    from scx_amd import accel
    assert accel.resolve_method("auto", 5000) == "randomized"
    assert accel.resolve_method("auto", 5001) == "randomized"
    assert accel.resolve_method("auto", 5000, "cpu") == "covariance"
    assert accel.resolve_method("auto", 5001, "cpu") == "randomized"
    assert accel.resolve_method("covariance", 30000) == "covariance"
    assert accel.resolve_method("randomized", 10) == "randomized"
    with pytest.raises(ValueError, match="Invalid method='pca'"):
        accel.resolve_method("pca", 10)
