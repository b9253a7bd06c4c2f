"""CPU oracle for the randomized sparse PCA path — test infrastructure only.

ctypes wrapper over ``oracle/liborc.so`` (built from ``scxpca_oracle.c`` by
``oracle/Makefile``). The C file restates scx-accel's CPU randomized PCA
(``scx-accel/src/pca/cpu.rs``) and the lazy normalize/log1p transform
(``pyscx/src/lazy_transform/transforms.rs``); each C function cites the
reference lines it follows.

Only ``tests/``, ``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of
``bench.py`` may import this package. The product (``scx_amd``) never does.

Parity status:
  * ChaCha8 stream, ``seed_from_u64``, ``gen_range``/``shuffle`` — pinned by the
    reference KAT ``seeded_permutation(16, 42)``.
  * ``next_u64`` pairing, ``Standard``/``Open01`` f64, ziggurat ``StandardNormal``
    — partially pinned (table invariants, published leading table entries,
    distribution tests); no reference fixture pins the sample bits.
  * X~ transform, column stats, SpMMs, QR, SVD, assembly, total variance —
    pinned against closed forms, brute-force dense f64 and numpy/LAPACK.
  * PFlog (``pflog_delta``, ``pflog_baseline``, ``estimate_alpha``,
    ``pflog_total_variance``, ``pflog_pca``) — pinned by the reference's own
    known answers (the hand-computed alpha cases, the dense v4 transform and
    its zero row sums, full-rank reconstructions of the centred / uncentred
    transform) and the three simulated-count fixtures with a known generating
    alpha; ``pflog_pca`` also against the dense SVD of the materialised
    ``Z = delta + b 1^T``, and the closed-form total variance against the
    dense ``Z``'s.
  * ``covariance_pca`` (numpy over the same X~, SpMM and total variance) —
    pinned against the dense SVD of the centred X~ (eigenvalues =
    sigma^2 / (n - 1), eigenvectors = right singular vectors up to sign,
    embeddings = U Sigma).
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liborc.so")
_lib = None

_i64 = ctypes.c_int64
_u64 = ctypes.c_uint64
_p = ctypes.c_void_p


class _Csr(ctypes.Structure):
    _fields_ = [("n_obs", _i64), ("n_vars", _i64), ("indptr", _p), ("indices", _p), ("data", _p)]


class _Params(ctypes.Structure):
    _fields_ = [
        ("n_comps", _i64),
        ("n_oversamples", _i64),
        ("n_power_iterations", _i64),
        ("zero_center", ctypes.c_int32),
        ("seed", _u64),
        ("target_sum", ctypes.c_double),
        ("log1p", ctypes.c_int32),
        ("nthreads", ctypes.c_int32),
    ]


def build(force: bool = False) -> str:
    """Compile liborc.so with the committed Makefile."""
    if force or not os.path.exists(_LIB_PATH):
        subprocess.check_call(["make", "-s", "-B" if force else "liborc.so"], cwd=_HERE)
    return _LIB_PATH


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        _lib = ctypes.CDLL(_LIB_PATH)
        _lib.orc_splitmix64.restype = _u64
        _lib.orc_splitmix64.argtypes = [_u64]
        _lib.orc_total_variance.restype = ctypes.c_double
        _lib.orc_randomized_pca.restype = ctypes.c_int
        _lib.orc_thin_svd.restype = ctypes.c_int
        _lib.orc_pflog_total_variance.restype = ctypes.c_double
        _lib.orc_pflog_pca.restype = ctypes.c_int
    return _lib


def _ptr(a: np.ndarray):
    return ctypes.c_void_p(a.ctypes.data)


class CsrView:
    """Keeps the numpy buffers alive for a ctypes ``orc_csr``."""

    def __init__(self, indptr, indices, data, shape):
        self.indptr = np.ascontiguousarray(indptr, dtype=np.int64)
        self.indices = np.ascontiguousarray(indices, dtype=np.uint32)
        self.data = np.ascontiguousarray(data, dtype=np.float32)
        self.shape = (int(shape[0]), int(shape[1]))
        self.c = _Csr(self.shape[0], self.shape[1], _ptr(self.indptr), _ptr(self.indices), _ptr(self.data))

    @classmethod
    def from_scipy(cls, m):
        return cls(m.indptr, m.indices, m.data, m.shape)

    def with_data(self, data):
        return CsrView(self.indptr, self.indices, data, self.shape)


def _as_view(x) -> CsrView:
    return x if isinstance(x, CsrView) else CsrView.from_scipy(x)


# ---------------------------------------------------------------- RNG
def splitmix64(x: int) -> int:
    return int(lib().orc_splitmix64(_u64(x)))


def seeded_permutation(n: int, seed: int) -> np.ndarray:
    out = np.empty(n, dtype=np.uint64)
    lib().orc_seeded_permutation(_u64(n), _u64(seed), _ptr(out))
    return out


def chacha8_block(key_words, counter: int = 0, nonce: int = 0) -> np.ndarray:
    key = np.ascontiguousarray(key_words, dtype=np.uint32)
    out = np.empty(16, dtype=np.uint32)
    lib().orc_chacha8_block(_ptr(key), _u64(counter), _u64(nonce), _ptr(out))
    return out


def chacha8_u32(seed: int, n: int) -> np.ndarray:
    out = np.empty(n, dtype=np.uint32)
    lib().orc_chacha8_u32_stream(_u64(seed), _u64(n), _ptr(out))
    return out


def chacha8_u64(seed: int, n: int) -> np.ndarray:
    out = np.empty(n, dtype=np.uint64)
    lib().orc_chacha8_u64_stream(_u64(seed), _u64(n), _ptr(out))
    return out


def zig_norm_tables():
    x = np.empty(257)
    f = np.empty(257)
    lib().orc_zig_norm_tables(_ptr(x), _ptr(f))
    return x, f


def random_gaussian(rows: int, cols: int, seed: int) -> np.ndarray:
    out = np.empty((rows, cols), dtype=np.float64)
    lib().orc_random_gaussian(_u64(rows), _u64(cols), _u64(seed), _ptr(out))
    return out


# ---------------------------------------------------------------- X~ / stats
def row_sums(x) -> np.ndarray:
    v = _as_view(x)
    out = np.empty(v.shape[0])
    lib().orc_row_sums(ctypes.byref(v.c), _ptr(out))
    return out


def transform(x, target_sum: float | None, log1p: bool) -> np.ndarray:
    v = _as_view(x)
    out = np.empty_like(v.data)
    lib().orc_transform(ctypes.byref(v.c), ctypes.c_double(target_sum or 0.0), ctypes.c_int(int(log1p)), _ptr(out))
    return out


def col_sums_and_sum_sq(x):
    v = _as_view(x)
    s = np.empty(v.shape[1])
    ss = np.empty(v.shape[1])
    lib().orc_col_sums_and_sum_sq(ctypes.byref(v.c), _ptr(s), _ptr(ss))
    return s, ss


# ---------------------------------------------------------------- SpMM / dense
def spmm_forward(x, m: np.ndarray, means=None, nthreads: int = 0) -> np.ndarray:
    v = _as_view(x)
    m = np.ascontiguousarray(m, dtype=np.float64)
    k = m.shape[1]
    y = np.empty((v.shape[0], k))
    mu = None if means is None else np.ascontiguousarray(means, dtype=np.float64)
    lib().orc_spmm_forward(ctypes.byref(v.c), _ptr(m), _i64(k), None if mu is None else _ptr(mu), _ptr(y), ctypes.c_int(nthreads))
    return y


def spmm_transpose(x, q: np.ndarray, means=None, nthreads: int = 0) -> np.ndarray:
    v = _as_view(x)
    q = np.ascontiguousarray(q, dtype=np.float64)
    k = q.shape[1]
    z = np.empty((v.shape[1], k))
    mu = None if means is None else np.ascontiguousarray(means, dtype=np.float64)
    lib().orc_spmm_transpose(ctypes.byref(v.c), _ptr(q), _i64(k), None if mu is None else _ptr(mu), _ptr(z), ctypes.c_int(nthreads))
    return z


def qr_thin_q(a: np.ndarray, nthreads: int = 0) -> np.ndarray:
    a = np.ascontiguousarray(a, dtype=np.float64)
    q = np.empty_like(a)
    lib().orc_qr_thin_q(_ptr(a), _i64(a.shape[0]), _i64(a.shape[1]), _ptr(q), ctypes.c_int(nthreads))
    return q


def thin_svd(b: np.ndarray):
    b = np.ascontiguousarray(b, dtype=np.float64)
    m, n = b.shape
    u = np.empty((m, n))
    s = np.empty(n)
    vt = np.empty((n, n))
    rc = lib().orc_thin_svd(_ptr(b), _i64(m), _i64(n), _ptr(u), _ptr(s), _ptr(vt))
    if rc != 0:
        raise RuntimeError("oracle Jacobi SVD did not converge")
    return u, s, vt


def total_variance(xt, col_sum_sq, means=None):
    v = _as_view(xt)
    css = np.ascontiguousarray(col_sum_sq, dtype=np.float64)
    mu = None if means is None else np.ascontiguousarray(means, dtype=np.float64)
    rec = ctypes.c_int(0)
    tv = lib().orc_total_variance(ctypes.byref(v.c), _ptr(css), None if mu is None else _ptr(mu), ctypes.byref(rec))
    return float(tv), bool(rec.value)


_PCA_ERRORS = {
    1: "n_components must be > 0",
    2: "matrix must be non-empty",
    3: "n_components ({nc}) exceeds matrix rank bound min(n_obs, n_vars) = {mn}",
    4: "SVD failed",
}


def randomized_pca(x, n_comps=50, n_oversamples=10, n_power_iterations=2, zero_center=True,
                   seed=0, target_sum=None, log1p=False, nthreads=0):
    """``randomized_pca_inmemory`` on the raw CSR ``x`` with the lazy transform
    (normalize_total to ``target_sum`` when given, then log1p) applied once.
    Returns a dict of f64 arrays."""
    v = _as_view(x)
    n, nv = v.shape
    k = max(1, min(n_comps + n_oversamples, nv, n))
    emb = np.zeros((n, max(n_comps, 1)))
    comps = np.zeros((max(n_comps, 1), nv))
    var = np.zeros(max(n_comps, 1))
    ratio = np.zeros(max(n_comps, 1))
    mean = np.zeros(max(nv, 1))
    sig = np.zeros(k)
    prm = _Params(n_comps, n_oversamples, n_power_iterations, int(bool(zero_center)), seed,
                  float(target_sum or 0.0), int(bool(log1p)), int(nthreads))
    rc = lib().orc_randomized_pca(ctypes.byref(v.c), ctypes.byref(prm), _ptr(emb), _ptr(comps), _ptr(var),
                                  _ptr(ratio), _ptr(mean), _ptr(sig))
    if rc != 0:
        raise ValueError(_PCA_ERRORS[rc].format(nc=n_comps, mn=min(n, nv)))
    return {
        "embeddings": emb,
        "components": comps,
        "variance": var,
        "variance_ratio": ratio,
        "mean": mean if zero_center else None,
        "sigma_all": sig,
    }


def covariance_pca(x, n_comps=50, zero_center=True, target_sum=None, log1p=False, nthreads=0):
    """``covariance_pca`` (scx-accel/src/pca/cpu.rs:1663-1854) on the raw CSR
    ``x`` with the lazy transform applied once: the f64 cross-product X~^T X~
    and column sums by ``orc_covariance_gram`` (the C restatement of
    accumulate_covariance_block, :198-291, rows in ascending order per entry),
    ``finalize_covariance`` (:303-322), the
    symmetric eigendecomposition (LAPACK syevd through numpy, as faer's
    self-adjoint EVD at :1748-1751), the top ``n_comps`` eigenpairs largest
    first (:1776-1785, eigenvalues clamped at 0), the guarded total variance
    (:1762-1771, :1029-1092) and the embeddings (X~ - mu) V by the forward
    SpMM (:1815-1846). Eigenvector signs are the solver's (the reference's
    are faer's): compare up to sign."""
    import scipy.sparse as sp

    v = _as_view(x)
    n, nv = v.shape
    if n_comps <= 0:
        raise ValueError(_PCA_ERRORS[1])
    if n == 0 or nv == 0:
        raise ValueError(_PCA_ERRORS[2])
    if n_comps > min(n, nv):
        raise ValueError(_PCA_ERRORS[3].format(nc=n_comps, mn=min(n, nv)))
    xt = v.with_data(transform(v, target_sum, log1p))
    g = np.empty((nv, nv))
    col_sums = np.empty(nv)
    if lib().orc_covariance_gram(ctypes.byref(xt.c), _ptr(g), _ptr(col_sums), int(nthreads)) != 0:
        raise ValueError("covariance oracle needs strictly increasing row indices")
    g = np.tril(g.T)  # column-major lower triangle -> row-major lower
    g = g + np.tril(g, -1).T
    means = col_sums / n if zero_center else None
    col_sum_sq = np.diag(g).copy()
    denom = max(n - 1.0, 1.0)
    cov = (g - n * np.outer(means, means)) / denom if zero_center else g / denom
    w, z = np.linalg.eigh(cov)
    tv, _ = total_variance(xt, col_sum_sq, means)
    idx = np.arange(nv - 1, nv - 1 - n_comps, -1)
    var = np.maximum(w[idx], 0.0)
    vmat = z[:, idx]
    ratio = var / tv if tv > 0 else np.zeros(n_comps)
    emb = spmm_forward(xt, np.ascontiguousarray(vmat), means, nthreads=nthreads)
    return {
        "embeddings": emb,
        "components": np.ascontiguousarray(vmat.T),
        "variance": var,
        "variance_ratio": ratio,
        "mean": means,
        "total_variance": tv,
    }


# ---------------------------------------------------------------- PFlog (v4)
ALPHA_MU_MIN = 1e-3           # AlphaOptions::default (scx-accel/src/pflog.rs:255-262)
ALPHA_FALLBACK = 0.25


def pflog_delta(x, four_alpha: float, form: int = 0) -> np.ndarray:
    """delta values on x's pattern: form 0 the lazy Scale{4a} -> Log1p chain,
    form 1 the in-memory builder (``orc_pflog_delta``)."""
    v = _as_view(x)
    out = np.empty_like(v.data)
    lib().orc_pflog_delta(ctypes.byref(v.c), ctypes.c_double(four_alpha), ctypes.c_int(int(form)), _ptr(out))
    return out


def pflog_baseline(delta) -> np.ndarray:
    v = _as_view(delta)
    out = np.empty(v.shape[0])
    lib().orc_pflog_baseline(ctypes.byref(v.c), _ptr(out))
    return out


def pflog_total_variance(col_sum_sq, colmean_delta, baseline, n_obs: int, n_vars: int) -> float:
    css = np.ascontiguousarray(col_sum_sq, dtype=np.float64)
    b = np.ascontiguousarray(baseline, dtype=np.float64)
    cd = None if colmean_delta is None else np.ascontiguousarray(colmean_delta, dtype=np.float64)
    return float(lib().orc_pflog_total_variance(_ptr(css), None if cd is None else _ptr(cd), _ptr(b),
                                                _i64(n_obs), _i64(n_vars)))


def column_moments(x):
    """streaming_mean_var (scx-accel/src/hvg/cpu.rs:107-130) on raw values:
    f64 column sums and sums of squares in nnz order, then
    finalize_column_moments (scx-sparse/src/moments.rs:193-233): mean = s / n,
    var = (ss - n * mean * mean) / max(n - 1, 1), negative clamped to 0."""
    v = _as_view(x)
    n, nv = v.shape
    if n == 0:
        return np.zeros(nv), np.zeros(nv)
    if not np.all(np.isfinite(v.data)):
        raise ValueError("non-finite value in the input")
    s, ss = col_sums_and_sum_sq(v)
    nf = float(n)
    means = s / nf
    var = (ss - nf * means * means) / max(nf - 1.0, 1.0)
    var = np.where(var < 0.0, 0.0, var)
    return means, var


def estimate_alpha(x, mu_min: float = ALPHA_MU_MIN, fallback_alpha: float = ALPHA_FALLBACK) -> dict:
    """estimate_alpha (scx-accel/src/pflog.rs:340-398): per-gene method-of-
    moments alpha_g = (var - mean) / mean^2 over every gene with mean > mu_min
    (finite ones), pooled by the median (the middle pair averaged for an even
    count, :403-415); the fallback alpha when the pool is empty or its median
    is not positive."""
    v = _as_view(x)
    if v.shape[1] == 0:
        raise ValueError("estimate_alpha requires n_vars > 0")
    means, var = column_moments(v)
    keep = means > mu_min
    m = means[keep]
    cand = (var[keep] - m) / (m * m)
    cand = np.sort(cand[np.isfinite(cand)])
    n_used = int(cand.size)
    fb = {"alpha": float(fallback_alpha), "pseudocount": 1.0 / (4.0 * fallback_alpha), "n_genes_used": n_used,
          "fell_back": True}
    if n_used == 0:
        return fb
    h = n_used // 2
    alpha = float(cand[h]) if n_used % 2 else float(0.5 * (cand[h - 1] + cand[h]))
    if not np.isfinite(alpha) or alpha <= 0.0:
        return fb
    return {"alpha": alpha, "pseudocount": 1.0 / (4.0 * alpha), "n_genes_used": n_used, "fell_back": False}


def pflog_pca(x, n_comps=50, n_oversamples=10, n_power_iterations=2, zero_center=True, seed=0, alpha=None,
              form=0, nthreads=0):
    """pflog_pca (scx-accel/src/pca/cpu.rs:815-866) on the raw CSR ``x``:
    alpha estimated (``alpha=None``) or pinned, delta = log1p(4 alpha x) in
    the given form, the baseline from delta (pflog.rs:186-222), then the
    C restatement. Returns a dict of f64 arrays plus the alpha fit, the
    baseline and the total variance."""
    v = _as_view(x)
    n, nv = v.shape
    if alpha is None:
        fit = estimate_alpha(v)
        fit["alpha_source"] = "estimated"
    else:
        fit = {"alpha": float(alpha), "pseudocount": 1.0 / (4.0 * alpha), "alpha_source": "pinned"}
    delta = v.with_data(pflog_delta(v, 4.0 * fit["alpha"], form))
    base = pflog_baseline(delta)
    k = max(1, min(n_comps + n_oversamples, nv, n))
    emb = np.zeros((n, max(n_comps, 1)))
    comps = np.zeros((max(n_comps, 1), nv))
    var = np.zeros(max(n_comps, 1))
    ratio = np.zeros(max(n_comps, 1))
    mean = np.zeros(max(nv, 1))
    sig = np.zeros(k)
    tv = ctypes.c_double(0.0)
    prm = _Params(n_comps, n_oversamples, n_power_iterations, int(bool(zero_center)), seed, 0.0, 0, int(nthreads))
    rc = lib().orc_pflog_pca(ctypes.byref(delta.c), _ptr(base), ctypes.byref(prm), _ptr(emb), _ptr(comps),
                             _ptr(var), _ptr(ratio), _ptr(mean), _ptr(sig), ctypes.byref(tv))
    if rc != 0:
        raise ValueError(_PCA_ERRORS[rc].format(nc=n_comps, mn=min(n, nv)))
    return {
        "embeddings": emb,
        "components": comps,
        "variance": var,
        "variance_ratio": ratio,
        "mean": mean if zero_center else None,
        "sigma_all": sig,
        "baseline": base,
        "delta": delta.data,
        "total_variance": tv.value,
        **fit,
    }
