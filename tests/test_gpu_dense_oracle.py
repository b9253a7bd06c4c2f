"""Results against a dense FP64 oracle: the exact PCA of the explicitly centred
X~ (LAPACK eigendecomposition of its Gram matrix, numpy f64), for the
10,000 x 2,000 workload in full and a 50,000-cell subsample of the
1,000,000 x 2,000 workload in the FP32 and BF16-sketch forms (the 20,000-cell
subsample of the 2,000,000 x 33,538 workload is in test_gpu_cfg5.py).

The randomized PCA approximates the exact one; its own error (what the f64
oracle of the reference CPU path makes on the same Omega) is reported beside
the GPU's. Asserted:
* Rayleigh-Ritz bound: every GPU singular value is at most the exact one
  (a Ritz value from a subspace never exceeds it; 2e-6 relative for f32);
* the GPU adds at most 1e-4 relative to the algorithm's own singular-value
  error (|s_gpu - s_exact| <= |s_f64 - s_exact| + 1e-4 s_exact);
* on the leading block a 2 % singular-value gap separates: per-PC embedding
  |cos| >= 0.99 against the exact PCA (the reference's GPU gate,
  pyscx/tests/test_accel_pca_gpu.py:176-181, there against the CPU route),
  and a principal angle to the exact block at most 1e-4 rad (FP32) or 1e-3
  rad (the BF16 sketch, the path's angle contract) above the f64 randomized
  PCA's own;
* variance-ratio drift against the f64 randomized PCA (the CPU path) < 1e-3
  (the same gate's second check); the drift against the exact ratios is
  reported;
* BF16 sketch against FP32: singular values <= 1e-4 relative.
"""
from __future__ import annotations

import json
import os
import sys

import numpy as np
import pytest

import oracle
from scx_amd import synth
from tests.pca_metrics import (dense_xt, exact_pca, max_principal_angle, per_pc_cosine, separated_rank,
                               sigma_rel_err)

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def record(label, row):
    """Adds a result line to gpurun_out/dense_oracle.jsonl when that folder exists."""
    if os.path.isdir("gpurun_out"):
        with open(os.path.join("gpurun_out", "dense_oracle.jsonl"), "a") as f:
            f.write(json.dumps({"case": label, **row}) + "\n")


def compare_exact(label, got, ex, ref64, nc, n_obs, total_var_exact, angle_slack=1e-4):
    """The assertions and the report line of one result against the exact PCA."""
    k = len(got["sigma"])
    s_ex = ex["sigma"][:k]
    e_gpu = sigma_rel_err(s_ex, got["sigma"][:k])
    e_f64 = sigma_rel_err(s_ex, ref64["sigma_all"][:k]) if ref64 is not None else None
    j = max(1, separated_rank(s_ex, 10, 0.02))
    ang = max_principal_angle(ex["components"][:j].T, got["components"][:j].T.astype(np.float64))
    cos = per_pc_cosine(ex["embeddings"][:, :j], got["embeddings"][:, :j])
    ratio_ex = s_ex[:nc] ** 2 / (n_obs - 1) / total_var_exact
    drift_ex = np.abs(got["variance_ratio"][:nc] - ratio_ex).max()
    row = {"sigma_err_top1": float(e_gpu[0]), "sigma_err_top10": float(e_gpu[:10].max()),
           "sigma_err_ncomps": float(e_gpu[:nc].max()), "sep_block": int(j), "angle_sep_block": float(ang),
           "cos_min_sep_block": float(cos.min()), "variance_ratio_drift_vs_exact": float(drift_ex)}
    if e_f64 is not None:
        ang64 = max_principal_angle(ex["components"][:j].T, ref64["components"][:j].T)
        drift64 = np.abs(got["variance_ratio"][:nc] - ref64["variance_ratio"][:nc]).max()
        row["f64_rpca_sigma_err_top10"] = float(e_f64[:10].max())
        row["f64_rpca_sigma_err_ncomps"] = float(e_f64[:nc].max())
        row["f64_rpca_angle_sep_block"] = float(ang64)
        row["gpu_vs_f64_rpca_sigma"] = float(sigma_rel_err(ref64["sigma_all"][:k], got["sigma"][:k]).max())
        row["gpu_vs_f64_rpca_ratio_drift"] = float(drift64)
    say(f"{label}: " + ", ".join(f"{a}={b:.3g}" if isinstance(b, float) else f"{a}={b}" for a, b in row.items()))
    record(label, row)
    assert np.all(got["sigma"][:k] <= s_ex * (1 + 2e-6)), "a Ritz value exceeds the exact singular value"
    assert cos.min() >= 0.99
    if e_f64 is not None:
        assert np.all(e_gpu <= e_f64 + 1e-4)
        assert ang <= ang64 + angle_slack
        assert drift64 < 1e-3
    return row


def _run_case(gpu, x, w, label, precisions, nthreads):
    n, nv = x.shape
    view = oracle.CsrView(x.indptr, x.indices, x.data, x.shape)
    say(f"case: {label} oracle f64 rPCA")
    ref = oracle.randomized_pca(view, n_comps=w.n_comps, n_oversamples=w.n_oversamples,
                                n_power_iterations=w.n_power_iterations, seed=0, target_sum=w.target_sum, log1p=True,
                                nthreads=nthreads)
    say(f"case: {label} exact dense FP64 PCA")
    d = dense_xt(x, w.target_sum)
    k = w.n_comps + w.n_oversamples
    ex = exact_pca(d, k)
    tv = float(((d - d.mean(axis=0)) ** 2).sum() / (n - 1))
    del d
    m = gpu.Matrix(x.indptr, x.indices, x.data, x.shape, target_sum=w.target_sum, log1p=True)
    out = {}
    try:
        for prec in precisions:
            say(f"case: {label} GPU rPCA precision={prec}")
            out[prec] = m.pca(n_comps=w.n_comps, n_oversamples=w.n_oversamples,
                              n_power_iterations=w.n_power_iterations, seed=0, qr_method=1, precision=prec)
    finally:
        m.close()
    for prec, got in out.items():
        compare_exact(f"{label} {'bf16' if prec else 'fp32'}", got, ex, ref, w.n_comps, n, tv,
                      1e-3 if prec else 1e-4)
    return out


def _threads():
    return max(1, min(16, len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else 1))


def test_cfg1_full_vs_dense_fp64(gpu):
    """10,000 x 2,000, k=50, p=10, q=2: the whole matrix."""
    w = synth.CFG1
    x = synth.generate(w)
    _run_case(gpu, x, w, "cfg1 10000x2000", (0,), _threads())


def test_cfg2_subsample_vs_dense_fp64(gpu):
    """The first 50,000 cells of the 1,000,000 x 2,000 workload, k=50, p=10,
    q=3: FP32 (config 2's form) and the BF16 sketch (config 3's), against the
    exact PCA of the subsample and against each other."""
    w = synth.CFG2
    x = synth.generate(w, n_obs=50_000)
    out = _run_case(gpu, x, w, "cfg2 50000x2000 subsample", (0, 1), _threads())
    se = sigma_rel_err(out[0]["sigma"][: w.n_comps], out[1]["sigma"][: w.n_comps]).max()
    say(f"cfg2 subsample: bf16 vs fp32 sigma {se:.2e}")
    record("cfg2 50000x2000 subsample bf16 vs fp32", {"sigma_rel": float(se)})
    assert se <= 1e-4
