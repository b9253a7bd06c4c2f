"""GPU PFlog (scxpca_matrix_create_pflog + scxpca_matrix_pca: the reference's
pflog_pca, scx-accel/src/pca/cpu.rs:815-866, with estimate_alpha and the
baseline of scx-accel/src/pflog.rs) against the oracle on the same counts.

Tolerances: alpha, n_genes_used and the fallback flag are compared exactly
(integer counts: the f64 column sums are exact in any order, and the moments
and median are the same IEEE operations). The baseline to 1e-9 relative (f64
row sums in another order). Column statistics of delta to 5e-7 relative (f32
values, f64 sums). The offset SpMMs to 2e-5 of the largest output (f32
accumulation). The PCA as the randomized path's parity: singular values
<= 1e-4 relative, principal angle of the top components <= 1e-3 rad, the mean
to 1e-6, the variance ratio to 2e-4; the closed-form total variance to 1e-9
relative (the same f64 formula on the same statistics).
"""
from __future__ import annotations

import os
import sys

import numpy as np
import pytest
import scipy.sparse as sp

import oracle
from scx_amd import synth
from tests.pca_metrics import max_principal_angle, sigma_rel_err
from tests.test_gpu_shapes import MATRICES

pytestmark = pytest.mark.gpu

GOLDEN = os.path.join(os.path.dirname(__file__), "golden", "pflog_alpha_fixtures.npz")
F64, LAZY = 1, 0


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def cfg1_head(n=3000):
    x = synth.generate(synth.CFG1, n_obs=n, workers=1)
    return sp.csr_matrix((x.data, x.indices.astype(np.int64), x.indptr), shape=x.shape)


def shape_case(name):
    c = MATRICES[name]()
    return sp.csr_matrix((c.data, c.indices.astype(np.int64), c.indptr), shape=c.shape)


def with_empty_rows(x, rows):
    d = x.tolil()
    for r in rows:
        d[r, :] = 0
    out = sp.csr_matrix(d, dtype=np.float32)
    out.eliminate_zeros()
    return out


def make(gpu, x, alpha=None, form=F64):
    return gpu.Matrix(x.indptr.astype(np.int64), x.indices.astype(np.uint32), x.data.astype(np.float32), x.shape,
                      pflog={"alpha": alpha, "delta_form": form})


def view(x):
    return oracle.CsrView(x.indptr, x.indices, x.data, x.shape)


@pytest.mark.parametrize("name", ["cfg1_head", "A", "B", "C", "tiny"])
def test_alpha_estimate_matches_oracle_exactly(gpu, name):
    if name == "cfg1_head":
        x = cfg1_head()
    elif name == "tiny":
        x = shape_case("tiny")
    else:
        x = sp.csr_matrix(np.load(GOLDEN)[f"{name}_counts"].astype(np.float32))
    say(f"case: pflog alpha {name} {x.shape}")
    ref = oracle.estimate_alpha(view(x))
    m = make(gpu, x)
    try:
        fit = m.pflog_fit()
    finally:
        m.close()
    assert fit["alpha_source"] == "estimated"
    assert fit["alpha"] == ref["alpha"] and fit["pseudocount"] == ref["pseudocount"]
    assert fit["n_genes_used"] == ref["n_genes_used"] and fit["fell_back"] == ref["fell_back"]


def test_alpha_simulated_fixtures_on_gpu(gpu):
    # the reference's assertions on its fixtures (pflog_reference_tests.rs:104-235), GPU estimator
    g = np.load(GOLDEN)
    fits = {}
    for fx in "ABC":
        m = make(gpu, sp.csr_matrix(g[f"{fx}_counts"].astype(np.float32)))
        try:
            fits[fx] = m.pflog_fit(baseline=False)
        finally:
            m.close()
    a, b, c = fits["A"], fits["B"], fits["C"]
    true_a = float(g["A_ALPHA_TRUE"])
    assert a["n_genes_used"] == int(g["A_N_GENES_POOLED"]) and not a["fell_back"]
    assert abs(a["alpha"] - true_a) * float(g["A_SEPARATION"]) < abs(float(g["A_TRUNCATED_ALPHA"]) - true_a)
    assert abs(a["alpha"] - true_a) / true_a <= float(g["A_REL_BAR"])
    assert not b["fell_back"] and b["n_genes_used"] == int(g["B_N_GENES_POOLED"])
    assert abs(b["alpha"] - float(g["B_ALPHA_TRUE"])) / float(g["B_ALPHA_TRUE"]) <= float(g["B_REL_BAR"])
    assert c["fell_back"] and c["alpha"] == 0.25 and c["n_genes_used"] == int(g["C_N_GENES_POOLED"])


@pytest.mark.parametrize("form", [F64, LAZY])
@pytest.mark.parametrize("name", ["cfg1_head", "ragged_empty", "wide3k"])
def test_baseline_and_delta_stats(gpu, name, form):
    x = cfg1_head() if name == "cfg1_head" else (
        with_empty_rows(shape_case("ragged"), [0, 7, 100]) if name == "ragged_empty" else shape_case("wide3k"))
    say(f"case: pflog baseline {name} form={form} {x.shape}")
    m = make(gpu, x, alpha=0.3, form=form)
    try:
        fit = m.pflog_fit()
        s, ss, cnt, rf = m.stats()
    finally:
        m.close()
    assert fit["alpha_source"] == "pinned" and fit["alpha"] == 0.3
    delta = view(x).with_data(oracle.pflog_delta(view(x), 1.2, form))
    ref_b = oracle.pflog_baseline(delta)
    np.testing.assert_allclose(fit["baseline"], ref_b, rtol=1e-9, atol=1e-300)
    empty = np.diff(x.indptr) == 0
    assert np.all(fit["baseline"][empty] == 0.0)
    os_, oss = oracle.col_sums_and_sum_sq(delta)
    np.testing.assert_allclose(s, os_, rtol=5e-7, atol=1e-9)
    np.testing.assert_allclose(ss, oss, rtol=5e-7, atol=1e-9)
    np.testing.assert_array_equal(rf, np.full(x.shape[0], 1.2))


@pytest.mark.parametrize("transpose", [False, True])
@pytest.mark.parametrize("k,centered", [(60, True), (60, False), (100, True), (5, True)])
def test_offset_spmm(gpu, transpose, k, centered):
    # streaming_spmm_forward_offset / _transpose_offset (cpu.rs:617-687)
    x = cfg1_head()
    say(f"case: pflog spmm transpose={transpose} k={k} centered={centered}")
    m = make(gpu, x, alpha=0.5)
    try:
        base = m.pflog_fit()["baseline"]
        delta = view(x).with_data(oracle.pflog_delta(view(x), 2.0, F64))
        mu = oracle.col_sums_and_sum_sq(delta)[0] / x.shape[0] + base.sum() / x.shape[0] if centered else None
        rng = np.random.default_rng(k + 7 * transpose)
        if transpose:
            q = rng.standard_normal((x.shape[0], k)) / np.sqrt(x.shape[0])
            ref = oracle.spmm_transpose(delta, q, mu) + (base @ q)[None, :]
            got = m.spmm(q, transpose=True, means=mu)
        else:
            w = rng.standard_normal((x.shape[1], k))
            ref = oracle.spmm_forward(delta, w, mu) + base[:, None] * w.sum(axis=0)[None, :]
            got = m.spmm(w, transpose=False, means=mu)
    finally:
        m.close()
    assert np.abs(got - ref).max() / np.abs(ref).max() < 2e-5


PCA_CASES = {
    "cfg1_head_est_f64": (lambda: cfg1_head(), None, F64, 50, 10, 2, True),
    "cfg1_head_est_f64_q4": (lambda: cfg1_head(), None, F64, 50, 10, 4, True),
    "cfg1_head_uncentered": (lambda: cfg1_head(), None, F64, 50, 10, 2, False),
    "cfg1_head_lazy_pinned": (lambda: cfg1_head(), 0.2, LAZY, 40, 20, 3, True),
    "cfg1_head_lazy_est": (lambda: cfg1_head(), None, LAZY, 50, 10, 2, True),
    "ragged_empty": (lambda: with_empty_rows(shape_case("ragged"), [0, 7, 100]), None, F64, 20, 10, 3, True),
    "wide3k_k100": (lambda: shape_case("wide3k"), None, F64, 100, 20, 3, True),
    "tiny": (lambda: shape_case("tiny"), None, F64, 5, 3, 3, True),
}


@pytest.mark.parametrize("case", list(PCA_CASES))
def test_pflog_pca_matches_oracle(gpu, case):
    make_x, alpha, form, nc, p, q, centered = PCA_CASES[case]
    x = make_x()
    say(f"case: pflog pca {case} {x.shape} nnz={x.nnz}")
    ref = oracle.pflog_pca(view(x), n_comps=nc, n_oversamples=p, n_power_iterations=q, zero_center=centered,
                           seed=0, alpha=alpha, form=form)
    m = make(gpu, x, alpha=alpha, form=form)
    try:
        fit = m.pflog_fit()
        got = m.pca(n_comps=nc, n_oversamples=p, n_power_iterations=q, zero_center=centered, seed=0,
                    qr_method=1)
    finally:
        m.close()
    assert fit["alpha"] == ref["alpha"]
    top = min(nc, 10)
    se = sigma_rel_err(ref["sigma_all"][:top], got["sigma"][:top])
    ang = max_principal_angle(ref["components"].T[:, :3], got["components"].T[:, :3].astype(np.float64))
    say(f"{case}: sigma rel err top{top} {se.max():.2e}, angle top3 {ang:.2e}")
    assert se.max() <= 1e-4
    assert ang <= 1e-3
    assert abs(got["total_variance"] - ref["total_variance"]) <= 1e-9 * ref["total_variance"]
    np.testing.assert_allclose(got["variance_ratio"][:top], ref["variance_ratio"][:top], rtol=2e-4)
    if centered:
        np.testing.assert_allclose(got["mean"], ref["mean"], rtol=1e-6, atol=1e-12)
    else:
        assert got["mean"] is None


def test_pflog_bf16_sketch_after_refresh(gpu):
    """The BF16 sketch form on a PFlog matrix (its bf16 copies come from the
    delta values), around a refresh. The sketch rounds delta to bf16 (~4e-3
    relative), which at 3,000 cells moves the top-10 Rayleigh-Ritz sigma by up
    to a few 1e-4 (measured 2.4e-4): bar 1e-3; it rotates the top-3 subspace
    by a few 1e-3 rad (measured 5.3e-3): bar 2e-2. An FP32 call after the
    refresh must meet the FP32 bars (1e-4, 1e-3 rad), which pins the
    rewritten values."""
    x = cfg1_head()
    say("case: pflog bf16 sketch")
    ref = oracle.pflog_pca(view(x), n_comps=50, n_oversamples=10, n_power_iterations=2, seed=0, alpha=None, form=F64)
    m = make(gpu, x)
    try:
        for precision, bar, abar in ((gpu.PRECISION_BF16_SKETCH, 1e-3, 2e-2), (gpu.PRECISION_BF16_SKETCH, 1e-3, 2e-2),
                                     (gpu.PRECISION_FP32, 1e-4, 1e-3)):
            got = m.pca(n_comps=50, n_oversamples=10, n_power_iterations=2, seed=0, qr_method=1, precision=precision)
            se = sigma_rel_err(ref["sigma_all"][:10], got["sigma"][:10])
            say(f"pflog precision={precision}: sigma rel err top10 {se.max():.2e}")
            assert se.max() <= bar
            ang = max_principal_angle(ref["components"].T[:, :3], got["components"].T[:, :3].astype(np.float64))
            say(f"pflog precision={precision}: angle top3 {ang:.2e}")
            assert ang <= abar
            m.refresh()
    finally:
        m.close()


@pytest.mark.parametrize("zero_center,alpha", [(True, 1.0), (False, 1.0), (True, 0.5)])
def test_full_rank_pca_reconstructs_the_dense_transform(gpu, zero_center, alpha):
    # pflog_tests.rs:336-454 on the GPU: embeddings @ components = (centred) Z
    rows = np.array([[0, 1, 3, 0], [2, 0, 0, 5], [1, 1, 1, 1], [4, 2, 0, 1], [0, 0, 6, 2], [3, 3, 1, 0]],
                    dtype=np.float32)
    x = sp.csr_matrix(rows)
    m = make(gpu, x, alpha=alpha, form=LAZY)
    try:
        got = m.pca(n_comps=4, n_oversamples=0, n_power_iterations=4, zero_center=zero_center, seed=42)
    finally:
        m.close()
    logs = np.log1p(4 * alpha * rows.astype(np.float64))
    z = logs - logs.mean(axis=1, keepdims=True)
    expected = z - z.mean(axis=0) if zero_center else z
    rec = got["embeddings"].astype(np.float64) @ got["components"].astype(np.float64)
    np.testing.assert_allclose(rec, expected, rtol=0, atol=1e-4)


def test_refresh_reestimates_alpha_on_changed_device_buffers(gpu):
    import torch

    x = cfg1_head(2000)
    dev = torch.device("cuda:0")
    ip = torch.from_numpy(x.indptr.astype(np.int64)).to(dev)
    ix = torch.from_numpy(x.indices.astype(np.int32)).to(dev)
    dv = torch.from_numpy(x.data.astype(np.float32)).to(dev)
    torch.cuda.synchronize()
    say("case: pflog refresh")
    m = gpu.Matrix(ip, ix, dv, x.shape, pflog={"alpha": None, "delta_form": F64})
    try:
        before = m.pflog_fit()
        dv.mul_(3.0)
        torch.cuda.synchronize()
        m.refresh()
        after = m.pflog_fit()
    finally:
        m.close()
    x3 = sp.csr_matrix((x.data * 3, x.indices, x.indptr), shape=x.shape)
    assert before["alpha"] == oracle.estimate_alpha(view(x))["alpha"]
    ref = oracle.estimate_alpha(view(x3))
    assert after["alpha"] == ref["alpha"] and after["n_genes_used"] == ref["n_genes_used"]
    d3 = view(x3).with_data(oracle.pflog_delta(view(x3), 4 * ref["alpha"], F64))
    np.testing.assert_allclose(after["baseline"], oracle.pflog_baseline(d3), rtol=1e-9, atol=1e-300)


ERRORS = [
    # (rows, alpha, form, expected message)
    ([[1, 0, 2], [0, -2, 1]], 0.5, F64, r"^pflog: negative count -2 at cell 1; counts must be non-negative$"),
    ([[np.nan, 1, 0], [0, 2, 1]], 0.5, F64, r"^pflog: non-finite count NaN at cell 0; counts must be finite$"),
    ([[1, 2, 0], [0, np.inf, 1]], None, F64, r"^HVG input contains a non-finite value \(inf\) at nonzero index 2;"),
    ([[1, 2, 0], [0, -3, 1]], 0.5, LAZY, r"^PFlog input contains a non-finite value \(NaN\) at nonzero index 2;"),
]


@pytest.mark.parametrize("rows,alpha,form,msg", ERRORS)
def test_count_checks(gpu, rows, alpha, form, msg):
    x = sp.csr_matrix(np.array(rows, dtype=np.float32))
    say(f"case: pflog count check {msg[:30]}")
    with pytest.raises(ValueError, match=msg):
        make(gpu, x, alpha=alpha, form=form)


def test_lazy_form_keeps_small_negative_counts(gpu):
    # the lazy chain has no sign check: log1p(4a x) is finite for x > -1/(4a)
    x = sp.csr_matrix(np.array([[1, 2, 0], [0, -0.5, 1]], dtype=np.float32))
    m = make(gpu, x, alpha=0.25, form=LAZY)
    try:
        b = m.pflog_fit()["baseline"]
    finally:
        m.close()
    ref = oracle.pflog_baseline(view(x).with_data(oracle.pflog_delta(view(x), 1.0, LAZY)))
    np.testing.assert_allclose(b, ref, rtol=1e-12)


def test_invalid_arguments_and_routes(gpu):
    x = shape_case("tiny")
    for a in (-1.0, float("nan"), float("inf")):
        with pytest.raises(ValueError, match="alpha must be positive and finite"):
            make(gpu, x, alpha=a)
    with pytest.raises(ValueError, match="n_vars > 0"):
        make(gpu, sp.csr_matrix((3, 0), dtype=np.float32))
    m = make(gpu, x)
    try:
        with pytest.raises(ValueError, match="covariance route is not defined for a PFlog matrix"):
            m.pca_covariance(n_comps=5)
    finally:
        m.close()
    plain = gpu.Matrix(x.indptr.astype(np.int64), x.indices.astype(np.uint32), x.data, x.shape, target_sum=1e4,
                       log1p=True)
    try:
        with pytest.raises(ValueError, match="not a PFlog matrix"):
            plain.pflog_fit()
    finally:
        plain.close()


def test_accel_pflog_end_to_end(gpu):
    from scx_amd import accel
    from scx_amd.adata import AnnDataLite

    x = cfg1_head(2000)
    say("case: accel.pflog store=pca")
    a = AnnDataLite(x.copy())
    accel.pflog(a, n_components=20, device="gpu")
    ref = oracle.pflog_pca(view(x), n_comps=20, n_oversamples=10, n_power_iterations=2, seed=0, alpha=None,
                           form=F64)
    u = a.uns["pflog"]
    assert u["version"] == "v4" and u["alpha_source"] == "estimated" and u["alpha"] == ref["alpha"]
    assert u["n_genes_used"] == ref["n_genes_used"] and u["fell_back"] == ref["fell_back"]
    np.testing.assert_allclose(a.obs["pflog_baseline"], ref["baseline"], rtol=1e-9)
    emb = a.obsm["X_pflog_pca"]
    assert emb.shape == (2000, 20) and emb.dtype == np.float32
    sv = a.uns["X_pflog_pca_singular_values"]
    np.testing.assert_allclose(sv[:5], np.sqrt(ref["variance"][:5] * 1999), rtol=1e-4)
    assert a.uns["scx_accel"]["pflog"]["route"] == "gpu_csr"
    # store="baseline": no PCA, the same baseline; a pinned alpha is stamped as pinned
    b = AnnDataLite(x.copy())
    accel.pflog(b, alpha=ref["alpha"], store="baseline", device="gpu")
    assert "X_pflog_pca" not in b.obsm and b.uns["pflog"]["alpha_source"] == "pinned"
    np.testing.assert_allclose(b.obs["pflog_baseline"], ref["baseline"], rtol=1e-9)


def test_pflog_general_layout(gpu, monkeypatch):
    """The general layout (no split offsets: the row-parallel forward kernel
    with the baseline epilogue, stats from the tile-CSC), forced by a zero
    metadata cap, against the oracle."""
    monkeypatch.setenv("SCXPCA_SPLIT_MAX_BYTES", "0")
    x = shape_case("wide3k")
    say("case: pflog general layout wide3k")
    ref = oracle.pflog_pca(view(x), n_comps=30, n_oversamples=10, n_power_iterations=3, seed=0, alpha=None, form=F64)
    m = make(gpu, x)
    try:
        fit = m.pflog_fit()
        got = m.pca(n_comps=30, n_oversamples=10, n_power_iterations=3, seed=0, qr_method=1)
    finally:
        m.close()
    assert fit["alpha"] == ref["alpha"]
    np.testing.assert_allclose(fit["baseline"], ref["baseline"], rtol=1e-9, atol=1e-300)
    se = sigma_rel_err(ref["sigma_all"][:10], got["sigma"][:10])
    say(f"pflog general: sigma rel err top10 {se.max():.2e}")
    assert se.max() <= 1e-4
    assert max_principal_angle(ref["components"].T[:, :3], got["components"].T[:, :3].astype(np.float64)) <= 1e-3
    assert abs(got["total_variance"] - ref["total_variance"]) <= 1e-9 * ref["total_variance"]


def test_accel_pflog_dense(gpu):
    """store="all": the dense Z in adata.layers equals the reference's
    materialisation (f32(b_i) + delta_ij in f32, pyscx/src/accel/pflog.rs:506-533)
    bit for bit on the oracle's delta and the returned baseline; the
    in-memory guard raises after the PCA outputs are written."""
    from scx_amd import accel
    from scx_amd.adata import AnnDataLite

    x = cfg1_head(500)
    say("case: accel.pflog store=all")
    a = AnnDataLite(x.copy())
    accel.pflog(a, store="all", n_components=10, device="gpu", layer_out="Z")
    ref = oracle.pflog_pca(view(x), n_comps=10, n_oversamples=10, n_power_iterations=2, seed=0, alpha=None, form=F64)
    b32 = np.asarray(a.obs["pflog_baseline"], dtype=np.float64).astype(np.float32)
    expect = np.repeat(b32[:, None], x.shape[1], axis=1)
    rows = np.repeat(np.arange(x.shape[0]), np.diff(x.indptr))
    expect[rows, x.indices] += ref["delta"]
    assert a.layers["Z"].dtype == np.float32 and a.layers["Z"].shape == x.shape
    np.testing.assert_array_equal(a.layers["Z"], expect)
    assert "X_pflog_pca" in a.obsm
    g = AnnDataLite(x.copy())
    with pytest.raises(RuntimeError, match="over the dense_max_elems=1000 guard"):
        accel.pflog(g, store="all", n_components=10, device="gpu", dense_max_elems=1000)
    assert "X_pflog_pca" in g.obsm and "pflog" not in g.layers


def test_accel_pflog_streams_when_not_resident(gpu, monkeypatch):
    """accel.pflog takes the streamed PFlog route when the residency pre-flight
    declines: the same alpha, baseline and spectrum, stamped resident_csr = False."""
    from scx_amd import accel
    from scx_amd.adata import AnnDataLite

    x = cfg1_head(2000)
    out = {}
    for env in ("1", "0"):
        monkeypatch.setenv("SCX_GPU_PCA_RESIDENT", env)
        a = AnnDataLite(x.copy())
        say(f"case: accel.pflog SCX_GPU_PCA_RESIDENT={env}")
        accel.pflog(a, n_components=20, device="gpu")
        out[env] = a
    r, s_ = out["1"], out["0"]
    assert s_.uns["scx_accel"]["pflog"]["resident_csr"] is False and r.uns["scx_accel"]["pflog"]["resident_csr"] is True
    assert s_.uns["pflog"] == r.uns["pflog"]
    np.testing.assert_array_equal(np.asarray(s_.obs["pflog_baseline"]), np.asarray(r.obs["pflog_baseline"]))
    np.testing.assert_allclose(s_.uns["X_pflog_pca_singular_values"], r.uns["X_pflog_pca_singular_values"], rtol=1e-5)


@pytest.mark.parametrize("store", ["baseline", "dense", "all"])
def test_accel_pflog_store_modes_stream_when_not_resident(gpu, monkeypatch, store):
    """Every store mode takes the streamed route when the residency pre-flight
    declines: the fit and baseline from a streamed pass, the dense layer built
    shard by shard (shards of 256 rows here, so several) with the fitted alpha
    pinned, bit-identical to the resident route's, and the PCA spectrum of
    store="all" within the streamed-vs-resident tolerance."""
    from scx_amd import accel
    from scx_amd.adata import AnnDataLite

    x = cfg1_head(700)
    out = {}
    for env in ("1", "0"):
        monkeypatch.setenv("SCX_GPU_PCA_RESIDENT", env)
        monkeypatch.setattr(accel, "stream_shard_rows", lambda *a, **k: 256)
        a = AnnDataLite(x.copy())
        say(f"case: accel.pflog store={store} SCX_GPU_PCA_RESIDENT={env}")
        accel.pflog(a, store=store, n_components=10, device="gpu", layer_out="Z")
        out[env] = a
    r, s_ = out["1"], out["0"]
    assert s_.uns["scx_accel"]["pflog"]["resident_csr"] is False and r.uns["scx_accel"]["pflog"]["resident_csr"] is True
    assert s_.uns["pflog"] == r.uns["pflog"]
    np.testing.assert_array_equal(np.asarray(s_.obs["pflog_baseline"]), np.asarray(r.obs["pflog_baseline"]))
    if store in ("dense", "all"):
        assert s_.layers["Z"].dtype == np.float32
        np.testing.assert_array_equal(s_.layers["Z"], r.layers["Z"])
    else:
        assert "Z" not in s_.layers
    if store == "all":
        np.testing.assert_allclose(s_.uns["X_pflog_pca_singular_values"], r.uns["X_pflog_pca_singular_values"],
                                   rtol=1e-5)
    else:
        assert "X_pflog_pca" not in s_.obsm
    if store == "dense":  # the in-memory guard holds on the streamed route too
        g = AnnDataLite(x.copy())
        with pytest.raises(RuntimeError, match="over the dense_max_elems=1000 guard"):
            accel.pflog(g, store="dense", n_components=10, device="gpu", dense_max_elems=1000)
