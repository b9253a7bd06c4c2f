"""GPU streamed PCA (scxpca_randomized_pca_streamed: the StreamingPcaOperator
route, scx-gpu/src/linear_operator.rs:61-266, for a host CSR larger than HBM)
against the resident matrix and the oracle on the same inputs.

The streamed route runs the resident kernels per shard, so against the
resident result only reduction order differs (per-shard f64 sums of the
transposed partials, per-shard column statistics): singular values to 1e-5
relative, the mean to 1e-12, the variance ratio to 1e-5, the top-3 principal
angle to 1e-4 rad. Against the oracle the randomized path's bars (1e-4, 1e-3
rad). Shard sizes cover one shard, uneven shards, shards of one 256-row tile
and shards smaller than a tile.
"""
from __future__ import annotations

import sys

import numpy as np
import pytest
import scipy.sparse as sp

import oracle
from scx_amd import synth
from tests.pca_metrics import max_principal_angle, sigma_rel_err
from tests.test_gpu_shapes import MATRICES

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def cfg1_head(n=3000):
    x = synth.generate(synth.CFG1, n_obs=n, workers=1)
    return sp.csr_matrix((x.data, x.indices.astype(np.int64), x.indptr), shape=x.shape)


def shape_case(name):
    c = MATRICES[name]()
    return sp.csr_matrix((c.data, c.indices.astype(np.int64), c.indptr), shape=c.shape)


def arrays(x):
    return x.indptr.astype(np.int64), x.indices.astype(np.uint32), x.data.astype(np.float32), x.shape


CASES = {
    "cfg1_one_shard": (cfg1_head, 3000, 50, 10, 2, True, 1),
    "cfg1_uneven_700": (cfg1_head, 700, 50, 10, 2, True, 1),
    "cfg1_tile_256": (cfg1_head, 256, 50, 10, 3, True, 0),
    "cfg1_small_97": (cfg1_head, 97, 20, 10, 2, True, 1),
    "cfg1_uncentered_1000": (cfg1_head, 1000, 50, 10, 2, False, 1),
    "wide3k_k100_400": (lambda: shape_case("wide3k"), 400, 100, 20, 3, True, 1),
    "ragged_333": (lambda: shape_case("ragged"), 333, 20, 10, 3, True, 0),
    "tiny_7": (lambda: shape_case("tiny"), 7, 5, 3, 3, True, 1),
}


@pytest.mark.parametrize("case", list(CASES))
def test_streamed_matches_resident_and_oracle(gpu, case):
    make_x, shard, nc, p, q, centered, qr = CASES[case]
    x = make_x()
    ip, ix, dv, shape = arrays(x)
    say(f"case: streamed {case} {shape} nnz={len(dv)} shard_rows={shard}")
    got = gpu.randomized_pca_streamed(ip, ix, dv, shape, shard, n_comps=nc, n_oversamples=p, n_power_iterations=q,
                                      zero_center=centered, seed=0, qr_method=qr, target_sum=1e4, log1p=True)
    m = gpu.Matrix(ip, ix, dv, shape, target_sum=1e4, log1p=True)
    try:
        res = m.pca(n_comps=nc, n_oversamples=p, n_power_iterations=q, zero_center=centered, seed=0, qr_method=qr)
    finally:
        m.close()
    ref = oracle.randomized_pca(oracle.CsrView(ip, ix, dv, shape), n_comps=nc, n_oversamples=p, n_power_iterations=q,
                                zero_center=centered, seed=0, target_sum=1e4, log1p=True)
    top = min(nc, 10)
    s_res = sigma_rel_err(res["sigma"][:top], got["sigma"][:top]).max()
    s_ref = sigma_rel_err(ref["sigma_all"][:top], got["sigma"][:top]).max()
    ang = max_principal_angle(res["components"].T[:, :3].astype(np.float64), got["components"].T[:, :3].astype(np.float64))
    say(f"{case}: sigma vs resident {s_res:.2e}, vs oracle {s_ref:.2e}, angle vs resident {ang:.2e}")
    assert s_res <= 1e-5
    assert s_ref <= 1e-4
    assert ang <= 1e-4
    assert max_principal_angle(ref["components"].T[:, :3], got["components"].T[:, :3].astype(np.float64)) <= 1e-3
    np.testing.assert_allclose(got["variance_ratio"][:top], res["variance_ratio"][:top], rtol=1e-5)
    assert abs(got["total_variance"] - res["total_variance"]) <= 1e-12 * res["total_variance"]
    if centered:
        np.testing.assert_allclose(got["mean"], res["mean"], rtol=1e-12, atol=1e-300)
    else:
        assert got["mean"] is None
    assert got["embeddings"].shape == (shape[0], nc) and np.all(np.isfinite(got["embeddings"]))


@pytest.mark.parametrize("n,shard", [(1000, 1000), (1000, 130)])
def test_streamed_stable_total_variance(gpu, n, shard):
    """The reference's cancellation matrix (scx-accel/src/pca/cpu.rs:2201-2214):
    the closed form cancels, the guard recomputes over the stored entries
    shard by shard; the centred variance is 3n / (n - 1) exactly."""
    rows = np.arange(n)
    dense = np.zeros((n, 3), dtype=np.float32)
    dense[:, 0] = 1.0e6 + np.where(rows % 2 == 0, 1.0, -1.0)
    dense[:, 1] = rows % 5
    x = sp.csr_matrix(dense)
    ip, ix, dv, shape = arrays(x)
    say(f"case: streamed cancellation n={n} shard_rows={shard}")
    got = gpu.randomized_pca_streamed(ip, ix, dv, shape, shard, n_comps=2, n_oversamples=1, n_power_iterations=2,
                                      seed=0)
    m = gpu.Matrix(ip, ix, dv, shape)
    try:
        res = m.pca(n_comps=2, n_oversamples=1, n_power_iterations=2, seed=0)
    finally:
        m.close()
    exact = 3.0 * n / (n - 1)
    assert got["tv_recomputed"] and res["tv_recomputed"]
    assert abs(got["total_variance"] - exact) <= 1e-9 * exact
    assert abs(res["total_variance"] - exact) <= 1e-9 * exact


def test_streamed_rejections(gpu):
    x = shape_case("tiny")
    ip, ix, dv, shape = arrays(x)
    with pytest.raises(ValueError, match="shard_rows must be positive"):
        gpu.randomized_pca_streamed(ip, ix, dv, shape, 0, n_comps=5)
    with pytest.raises(ValueError, match="exceeds matrix rank bound"):
        gpu.randomized_pca_streamed(ip, ix, dv, shape, 10, n_comps=31)
    bad = ip.copy()
    bad[5] = bad[4] - 1
    with pytest.raises(ValueError, match="indptr must be non-decreasing"):
        gpu.randomized_pca_streamed(bad, ix, dv, shape, 10, n_comps=5)
    unsorted = ix.copy()
    a, b = ip[3], ip[4]
    assert b - a >= 2
    unsorted[a], unsorted[a + 1] = unsorted[a + 1], unsorted[a]
    with pytest.raises(ValueError, match="sorted column indices"):
        gpu.randomized_pca_streamed(ip, unsorted, dv, shape, 10, n_comps=5)


def test_streamed_from_page_locked_host_memory(gpu):
    """Page-locked host arrays are sent to the device directly (no staging
    copy); the result is bit-identical to the pageable (staged) run."""
    import torch

    x = cfg1_head(2000)
    ip, ix, dv, shape = arrays(x)
    pix = torch.from_numpy(ix.view(np.int32)).pin_memory()
    pdv = torch.from_numpy(dv).pin_memory()
    pinned_ix, pinned_dv = pix.numpy().view(np.uint32), pdv.numpy()
    assert pinned_ix.ctypes.data == pix.data_ptr() and pinned_dv.ctypes.data == pdv.data_ptr()
    say("case: streamed page-locked vs pageable")
    kw = dict(n_comps=20, n_oversamples=10, n_power_iterations=2, seed=0, qr_method=1, target_sum=1e4, log1p=True)
    a = gpu.randomized_pca_streamed(ip, ix, dv, shape, 300, **kw)
    b = gpu.randomized_pca_streamed(ip, pinned_ix, pinned_dv, shape, 300, **kw)
    np.testing.assert_array_equal(a["sigma"], b["sigma"])
    np.testing.assert_array_equal(a["embeddings"], b["embeddings"])
    np.testing.assert_array_equal(a["components"], b["components"])


def test_accel_pca_streams_when_not_resident(gpu, monkeypatch):
    """accel.pca takes the streamed route when the residency pre-flight
    declines (SCX_GPU_PCA_RESIDENT=0 forces it, as in
    scx-gpu/src/gpu_pca_resident.rs:71-74): same result as the resident route,
    stamped resident_csr = False; the BF16 sketch needs residency."""
    from scx_amd import accel
    from scx_amd.adata import AnnDataLite

    x = cfg1_head(2000)
    out = {}
    for env in ("1", "0"):
        monkeypatch.setenv("SCX_GPU_PCA_RESIDENT", env)
        a = AnnDataLite(x.copy())
        accel.normalize_total(a, target_sum=1e4)
        accel.log1p(a)
        say(f"case: accel.pca SCX_GPU_PCA_RESIDENT={env}")
        accel.pca(a, n_comps=20, device="gpu", method="randomized")
        out[env] = a
    r, s_ = out["1"], out["0"]
    assert r.uns["scx_accel"]["pca"]["resident_csr"] is True
    assert s_.uns["scx_accel"]["pca"]["resident_csr"] is False
    np.testing.assert_allclose(s_.uns["pca"]["variance"], r.uns["pca"]["variance"], rtol=1e-5)
    assert max_principal_angle(r.varm["PCs"][:, :3].astype(np.float64), s_.varm["PCs"][:, :3].astype(np.float64)) <= 1e-4
    a = AnnDataLite(x.copy())
    with pytest.raises(RuntimeError, match="needs the matrix resident"):
        accel.pca(a, n_comps=20, device="gpu", method="randomized", precision="bf16")


def test_resident_build_failure_declines_to_streaming(gpu, monkeypatch):
    """A runtime failure while the resident matrix is built (forced from the
    host by SCXPCA_FAIL_RESIDENT_BUILD=1, reported as out of memory) is a
    decline: accel.pca and the one-shot C entry continue on the streamed
    operator and return the resident route's result, stamped
    resident_csr = False (decline_on_runtime_failure,
    scx-gpu/src/gpu_pca.rs:547-550, scx-gpu/src/error.rs:141-162). An input
    defect still fails the call."""
    from scx_amd import accel
    from scx_amd.adata import AnnDataLite

    x = cfg1_head(2000)
    ip, ix, dv, shape = arrays(x)
    kw = dict(n_comps=20, n_oversamples=10, n_power_iterations=2, seed=0, qr_method=1, target_sum=1e4, log1p=True)
    say("case: one-shot resident")
    ref = gpu.randomized_pca(ip, ix, dv, shape, **kw)
    assert ref["resident"] is True
    monkeypatch.setenv("SCXPCA_FAIL_RESIDENT_BUILD", "1")
    with pytest.raises(MemoryError, match="forced by SCXPCA_FAIL_RESIDENT_BUILD"):
        gpu.Matrix(ip, ix, dv, shape, target_sum=1e4, log1p=True)
    say("case: one-shot with the resident build failing")
    got = gpu.randomized_pca(ip, ix, dv, shape, **kw)
    assert got["resident"] is False
    se = sigma_rel_err(ref["sigma"][:10], got["sigma"][:10]).max()
    say(f"declined one-shot: sigma vs resident {se:.2e}")
    assert se <= 1e-5
    a = AnnDataLite(x.copy())
    accel.normalize_total(a, target_sum=1e4)
    accel.log1p(a)
    say("case: accel.pca with the resident build failing")
    accel.pca(a, n_comps=20, device="gpu")
    assert a.uns["scx_accel"]["pca"]["resident_csr"] is False
    np.testing.assert_allclose(a.uns["pca"]["variance"][:10], ref["variance"][:10], rtol=1e-5)
    # an input defect still fails the call (here found by the streamed
    # operator's own checks); the BF16 sketch cannot stream, so its build
    # failure propagates
    bad = ix.copy()
    bad[5] = shape[1] + 3
    with pytest.raises(ValueError, match="column index out of range"):
        gpu.randomized_pca(ip, bad, dv, shape, **kw)
    with pytest.raises(MemoryError, match="forced by SCXPCA_FAIL_RESIDENT_BUILD"):
        gpu.randomized_pca(ip, ix, dv, shape, precision=gpu.PRECISION_BF16_SKETCH, **kw)
    monkeypatch.delenv("SCXPCA_FAIL_RESIDENT_BUILD")
    monkeypatch.setenv("SCX_GPU_PCA_RESIDENT", "0")
    say("case: one-shot with residency declined by the pre-flight")
    pre = gpu.randomized_pca(ip, ix, dv, shape, **kw)
    assert pre["resident"] is False
    np.testing.assert_array_equal(pre["sigma"], got["sigma"])


def test_streamed_error_positions_are_global(gpu):
    """A bad count in a later shard is reported at its position in the whole
    matrix (cell and nonzero index), as the resident route reports it."""
    x = cfg1_head(1200)
    ip, ix, dv, shape = arrays(x)
    p_bad = int(ip[900]) + 3
    cell = 900
    neg = dv.copy()
    neg[p_bad] = -2.0
    say("case: streamed pflog negative count in shard 4 of 300-row shards")
    with pytest.raises(ValueError, match=f"negative count -2 at cell {cell};"):
        gpu.pflog_pca_streamed(ip, ix, neg, shape, 300, alpha=0.2, delta_form=1, n_comps=10)
    with pytest.raises(ValueError, match=f"negative count -2 at cell {cell};"):
        gpu.Matrix(ip, ix, neg, shape, pflog={"alpha": 0.2, "delta_form": 1})
    nan = dv.copy()
    nan[p_bad] = np.inf
    say("case: streamed pflog non-finite value (lazy form) in shard 4")
    with pytest.raises(ValueError, match=f"at nonzero index {p_bad};"):
        gpu.pflog_pca_streamed(ip, ix, nan, shape, 300, alpha=0.2, delta_form=0, n_comps=10)
    with pytest.raises(ValueError, match=f"at nonzero index {p_bad};"):
        gpu.Matrix(ip, ix, nan, shape, pflog={"alpha": 0.2, "delta_form": 0})


def test_streamed_general_layout(gpu, monkeypatch):
    """Shards in the general layout (zero metadata cap: no split offsets, the
    row-parallel forward, stats from the tile-CSC) against the resident
    matrix in the same layout."""
    monkeypatch.setenv("SCXPCA_SPLIT_MAX_BYTES", "0")
    x = shape_case("wide3k")
    ip, ix, dv, shape = arrays(x)
    say("case: streamed general layout wide3k shard_rows=500")
    kw = dict(n_comps=30, n_oversamples=10, n_power_iterations=3, seed=0, qr_method=1)
    got = gpu.randomized_pca_streamed(ip, ix, dv, shape, 500, target_sum=1e4, log1p=True, **kw)
    m = gpu.Matrix(ip, ix, dv, shape, target_sum=1e4, log1p=True)
    try:
        res = m.pca(**kw)
    finally:
        m.close()
    se = sigma_rel_err(res["sigma"][:10], got["sigma"][:10]).max()
    say(f"streamed general: sigma vs resident {se:.2e}")
    assert se <= 1e-5
    np.testing.assert_allclose(got["mean"], res["mean"], rtol=1e-12, atol=1e-300)


@pytest.mark.parametrize("alpha,form,shard", [(None, 1, 700), (0.2, 0, 97), (None, 0, 3000)])
def test_streamed_pflog_matches_resident(gpu, alpha, form, shard):
    """PFlog over streamed shards (scxpca_pflog_pca_streamed) against the
    resident PFlog matrix: alpha and the baseline identical (the raw moments
    are exact sums; each row's baseline is summed in the same order), the
    spectrum to 1e-5, the closed-form total variance to 1e-12."""
    x = cfg1_head(3000)
    ip, ix, dv, shape = arrays(x)
    say(f"case: streamed pflog alpha={alpha} form={form} shard_rows={shard}")
    kw = dict(n_comps=30, n_oversamples=10, n_power_iterations=3, seed=0, qr_method=1)
    got, fit = gpu.pflog_pca_streamed(ip, ix, dv, shape, shard, alpha=alpha, delta_form=form, **kw)
    m = gpu.Matrix(ip, ix, dv, shape, pflog={"alpha": alpha, "delta_form": form})
    try:
        rfit = m.pflog_fit()
        res = m.pca(**kw)
    finally:
        m.close()
    assert fit["alpha"] == rfit["alpha"] and fit["alpha_source"] == rfit["alpha_source"]
    np.testing.assert_array_equal(fit["baseline"], rfit["baseline"])
    se = sigma_rel_err(res["sigma"][:10], got["sigma"][:10]).max()
    say(f"streamed pflog: sigma vs resident {se:.2e}")
    assert se <= 1e-5
    assert abs(got["total_variance"] - res["total_variance"]) <= 1e-12 * res["total_variance"]
    np.testing.assert_allclose(got["mean"], res["mean"], rtol=1e-12, atol=1e-300)
    assert max_principal_angle(res["components"].T[:, :3].astype(np.float64), got["components"].T[:, :3].astype(np.float64)) <= 1e-4
