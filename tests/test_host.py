"""Host-side tests (no GPU): the C ABI library loads and exports every declared
symbol, the pca(...) surface validates like the reference, the lazy transform
chain matches the oracle's X~, and the synthetic generator is reproducible."""
from __future__ import annotations

import os
import re

import numpy as np
import pytest
import scipy.sparse as sp

import oracle
from scx_amd import _lib, accel, synth
from scx_amd.adata import AnnDataLite

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_library_exports_every_declared_symbol():
    inc = os.path.join(ROOT, "include")
    header = "".join(open(os.path.join(inc, f)).read() for f in sorted(os.listdir(inc)) if f.endswith(".h"))
    declared = set(re.findall(r"\b(scx(?:pca|shard)_[a-z0-9_]+)\s*\(", header))
    assert declared, "no declarations parsed"
    assert declared == set(_lib.EXPORTS), declared ^ set(_lib.EXPORTS)
    lib = _lib.lib()
    for name in declared:
        assert hasattr(lib, name), name


def test_device_grammar():
    assert accel.resolve_device("cpu") == ("cpu", None)
    with pytest.raises(ValueError, match="unknown device: 'tpu'"):
        accel.resolve_device("tpu")
    with pytest.raises(ValueError, match="non-integer suffix"):
        accel.resolve_device("gpu:x")
    with pytest.raises(ValueError, match="unknown device"):
        accel.resolve_device("gpu0")
    if not accel.gpu_available():
        assert accel.resolve_device("auto") == ("cpu", None)
        with pytest.raises(RuntimeError, match="no ROCm GPU found"):
            accel.resolve_device("gpu")


def _adata(n=50, nv=20):
    c = synth.random_csr(n, nv, 0.3, seed=1)
    return AnnDataLite(c.to_scipy())


@pytest.mark.parametrize("kw,msg", [
    (dict(method="pca"), "Invalid method='pca'; expected 'auto', 'covariance', or 'randomized'"),
    (dict(qr_method="qr"), "Invalid qr_method='qr'; expected 'householder' or 'cholesky'"),
    (dict(spmm_policy="x"), "Invalid spmm_policy='x'"),
    (dict(prefer_format="csc"), "only supports prefer_format='csr'"),
])
def test_pca_argument_validation(kw, msg):
    with pytest.raises(ValueError, match=re.escape(msg)):
        accel.pca(_adata(), device="cpu", **kw)


def test_pca_cpu_route_is_not_silently_served():
    with pytest.raises(RuntimeError, match="GPU route of pca only"):
        accel.pca(_adata(), device="cpu")


@pytest.mark.parametrize("kw,exc,msg", [
    (dict(alpha=-1.0), ValueError, "pflog: alpha must be positive and finite, got -1.0"),
    (dict(alpha=float("nan")), ValueError, "pflog: alpha must be positive and finite"),
    (dict(store="x"), ValueError, 'pflog: unknown store \'x\'; expected "pca", "baseline", "dense", or "all"'),
    (dict(store_repr="csr"), ValueError, "pflog: unknown store_repr"),
    (dict(out="z.scx"), ValueError, "pflog: out= streams the materialized transform to disk"),
    (dict(layer="counts"), ValueError, "layer 'counts' not found in adata.layers"),
    (dict(store="dense", out="z.scx"), NotImplementedError, "pflog: out= writes an SCX file"),
    (dict(), RuntimeError, "GPU route of pflog only"),
])
def test_pflog_argument_validation(kw, exc, msg):
    # pyscx/src/accel/pflog.rs:113-174: checked in this order, before any device work
    with pytest.raises(exc, match=re.escape(msg)):
        accel.pflog(_adata(), device="cpu", **kw)


def test_residency_preflight(monkeypatch):
    # SCX_GPU_PCA_RESIDENT=0 declines before any device query (gpu_pca_resident.rs:71-74)
    monkeypatch.setenv("SCX_GPU_PCA_RESIDENT", "0")
    assert accel.resident_fits(1000, 100, 10_000, 60, 0) is False
    # the footprint model grows with every dimension and sits near 19-30 bytes per nonzero at 1M x 2k
    b = accel._resident_bytes(1_000_000, 2000, 162_701_712, 60)
    assert accel._resident_bytes(2_000_000, 2000, 162_701_712, 60) > b
    assert accel._resident_bytes(1_000_000, 4000, 162_701_712, 60) > b
    assert accel._resident_bytes(1_000_000, 2000, 200_000_000, 60) > b
    assert 19 * 162_701_712 < b < 30 * 162_701_712


def test_lazy_chain_matches_oracle_transform():
    c = synth.random_csr(200, 60, 0.2, seed=2)
    ad = AnnDataLite(c.to_scipy())
    accel.normalize_total(ad, target_sum=1e4, device="cpu")
    accel.log1p(ad, device="cpu")
    assert isinstance(ad.X, accel.LazyCsr) and ad.X.target_sum == 1e4 and ad.X.log1p
    got = ad.X.materialize().data
    ref = oracle.transform(oracle.CsrView(c.indptr, c.indices, c.data, c.shape), 1e4, True)
    # numpy float32 log1p vs glibc log1pf: within one f32 ulp
    assert np.all(np.abs(got - ref) <= np.spacing(np.abs(ref)))


def test_normalize_total_default_target_is_median_positive_total():
    m = sp.csr_matrix(np.array([[1.0, 1.0], [0.0, 0.0], [3.0, 1.0], [5.0, 5.0]], np.float32))
    ad = AnnDataLite(m)
    accel.normalize_total(ad, device="cpu")
    assert ad.X.target_sum == 4.0  # median of (2, 4, 10)


def test_mask_var_resolution():
    ad = _adata()
    ad.var["highly_variable"] = np.arange(20) % 2 == 0
    cols, name = accel._resolve_mask_var(ad, None, 20)
    assert name == "highly_variable" and cols.tolist() == list(range(0, 20, 2))
    with pytest.raises(ValueError, match="mask_var length 3 != n_vars 20"):
        accel._resolve_mask_var(ad, np.array([True, False, True]), 20)
    with pytest.raises(ValueError, match="all-false mask"):
        accel._resolve_mask_var(ad, np.zeros(20, bool), 20)
    with pytest.raises(ValueError, match="mask_var column 'nope' not found"):
        accel._resolve_mask_var(ad, "nope", 20)


def test_generator_reproducible_and_prefix_stable():
    a = synth.generate(synth.CFG1, n_obs=5000, workers=1)
    b = synth.generate(synth.CFG1, n_obs=5000, workers=2)
    np.testing.assert_array_equal(a.indptr, b.indptr)
    np.testing.assert_array_equal(a.indices, b.indices)
    np.testing.assert_array_equal(a.data, b.data)
    head = synth.generate(synth.CFG1, n_obs=3000, workers=1)
    np.testing.assert_array_equal(head.indices, a.indices[: head.nnz])
    # canonical rows and integer counts
    m = a.to_scipy()
    assert m.has_sorted_indices and np.all(a.data >= 1) and np.all(a.data == np.round(a.data))
    assert abs(a.nnz / (5000 * 2000) - 0.10) < 0.01
