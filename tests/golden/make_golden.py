"""Writes tests/golden/golden_v1.npz: the build's own golden vectors for this
path, produced by the oracle (the C restatement of scx-accel/src/pca/cpu.rs)
and checked against independent restatements by tests/test_golden.py.

Contents (all seeds fixed):
* xt_10x5: X~ = log1p(normalize_total(X, 1e4)) of the reference fixture
  test_csr_10x5 (scx-accel/src/pca/cpu.rs:2326), f32 in CSR order.
* omega_s{0,42,12345}: the first 4,096 values of the Omega stream per seed
  (ChaCha8Rng::seed_from_u64 + StandardNormal, cpu.rs:1174-1178), f64.
* spmm_*: a 200 x 50 CSR cut from the 10k x 2k workload generator (stored),
  panels M (50 x 7) and Q (200 x 7), and the oracle's forward / transposed
  SpMM of X~, plain and centered (cpu.rs:1185-1267), f64.
* pca_*: the oracle's randomized PCA (cpu.rs:535-605) of the first 2,000 rows
  and 300 genes of the 10k x 2k workload (regenerated by the tests; its
  SHA-256 is stored), k=50, p=10, q=2, seed 0: all singular values, the mean,
  variance and variance ratio, and the first 10 components and embeddings.

Run: python tests/golden/make_golden.py
"""
from __future__ import annotations

import hashlib
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

import oracle  # noqa: E402
from scx_amd import synth  # noqa: E402

OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden_v1.npz")

CSR_10X5 = dict(
    indptr=np.array([0, 2, 4, 7, 9, 12, 14, 17, 19, 22, 24], dtype=np.int64),
    indices=np.array([0, 2, 1, 3, 0, 2, 4, 1, 4, 0, 2, 3, 1, 3, 0, 2, 4, 1, 3, 0, 2, 4, 1, 3], dtype=np.uint32),
    data=np.array([1.0, 3.0, 2.0, 4.0, 5.0, 1.0, 2.0, 3.0, 6.0, 2.0, 4.0, 1.0, 1.0, 3.0, 3.0, 2.0,
                   5.0, 4.0, 2.0, 1.0, 5.0, 3.0, 2.0, 1.0], dtype=np.float32),
    shape=(10, 5),
)
OMEGA_SEEDS = (0, 42, 12345)
PCA_ROWS, PCA_COLS = 2000, 300


def slice_csr(x, n_rows, n_cols):
    """Rows [0, n_rows) and genes [0, n_cols) of a synth CSR, canonical."""
    m = x.to_scipy()[:n_rows, :n_cols].tocsr()
    m.sort_indices()
    return (m.indptr.astype(np.int64), m.indices.astype(np.uint32), m.data.astype(np.float32), m.shape)


def pca_input():
    base = synth.generate(synth.CFG1, n_obs=PCA_ROWS, workers=1)
    return slice_csr(base, PCA_ROWS, PCA_COLS)


def digest(indptr, indices, data) -> str:
    h = hashlib.sha256()
    for a in (indptr, indices, data):
        h.update(np.ascontiguousarray(a).tobytes())
    return h.hexdigest()


def main() -> None:
    g = {}
    v = oracle.CsrView(CSR_10X5["indptr"], CSR_10X5["indices"], CSR_10X5["data"], CSR_10X5["shape"])
    g["xt_10x5"] = oracle.transform(v, 1e4, True).astype(np.float32)
    for s in OMEGA_SEEDS:
        g[f"omega_s{s}"] = oracle.random_gaussian(64, 64, s).reshape(-1)

    base = synth.generate(synth.CFG1, n_obs=200, workers=1)
    ip, ix, dv, shp = slice_csr(base, 200, 50)
    sv = oracle.CsrView(ip, ix, dv, shp)
    xt = sv.with_data(oracle.transform(sv, 1e4, True))
    rng = np.random.default_rng(7)
    m = rng.standard_normal((50, 7))
    q = rng.standard_normal((200, 7))
    sums, _ = oracle.col_sums_and_sum_sq(xt)
    mu = sums / shp[0]
    g.update(spmm_indptr=ip, spmm_indices=ix, spmm_data=dv, spmm_m=m, spmm_q=q, spmm_mu=mu,
             spmm_fwd=oracle.spmm_forward(xt, m), spmm_fwd_centered=oracle.spmm_forward(xt, m, mu),
             spmm_t=oracle.spmm_transpose(xt, q), spmm_t_centered=oracle.spmm_transpose(xt, q, mu))

    ip, ix, dv, shp = pca_input()
    res = oracle.randomized_pca(oracle.CsrView(ip, ix, dv, shp), n_comps=50, n_oversamples=10,
                                n_power_iterations=2, seed=0, target_sum=1e4, log1p=True)
    g.update(pca_input_sha256=np.array(digest(ip, ix, dv)), pca_sigma=res["sigma_all"], pca_mean=res["mean"],
             pca_variance=res["variance"], pca_variance_ratio=res["variance_ratio"],
             pca_components10=res["components"][:10], pca_embeddings10=res["embeddings"][:, :10])
    np.savez_compressed(OUT, **g)
    print(f"wrote {OUT} ({os.path.getsize(OUT)} bytes)")


if __name__ == "__main__":
    main()
