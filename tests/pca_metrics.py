"""Comparison metrics for PCA results (shared by the CPU and GPU tests)."""
from __future__ import annotations

import numpy as np


def max_principal_angle(a: np.ndarray, b: np.ndarray) -> float:
    """Largest principal angle (radians) between the column spans of a and b."""
    qa, _ = np.linalg.qr(np.asarray(a, dtype=np.float64))
    qb, _ = np.linalg.qr(np.asarray(b, dtype=np.float64))
    # sin(theta_max) = ||(I - Qa Qa^T) Qb||_2 (accurate for small angles,
    # unlike arccos of the cosines)
    resid = qb - qa @ (qa.T @ qb)
    s = np.linalg.norm(resid, 2)
    return float(np.arcsin(min(1.0, s)))


def separated_rank(sigma, max_j: int = 10, gap: float = 0.02) -> int:
    """Largest j <= max_j whose trailing relative gap (s_j - s_{j+1}) / s_j is at
    least `gap`: the top-j subspace is then well defined (sin-theta bound), and
    principal angles against it are meaningful. 0 when no such j exists."""
    s = np.asarray(sigma, dtype=np.float64)
    best = 0
    for j in range(1, min(max_j, len(s) - 1) + 1):
        if s[j - 1] > 0 and (s[j - 1] - s[j]) / s[j - 1] >= gap:
            best = j
    return best


def align_signs(ref: np.ndarray, x: np.ndarray, axis: int = 0) -> np.ndarray:
    """Flip components of x (columns when axis=0) to correlate positively with ref."""
    dots = np.sum(ref * x, axis=axis)
    sgn = np.where(dots < 0, -1.0, 1.0)
    return x * (sgn if axis == 0 else sgn[:, None])


def score_rel_err(ref: np.ndarray, x: np.ndarray) -> np.ndarray:
    """Per-component relative L2 error of embeddings (columns) after sign alignment."""
    xa = align_signs(ref, x, axis=0)
    return np.linalg.norm(xa - ref, axis=0) / np.maximum(np.linalg.norm(ref, axis=0), 1e-300)


def sigma_rel_err(ref: np.ndarray, x: np.ndarray) -> np.ndarray:
    return np.abs(np.asarray(x) - np.asarray(ref)) / np.maximum(np.abs(ref), 1e-300)


def dense_xtilde(csr, target_sum, log1p):
    """Explicit dense X~ (f64) of a small CSR via the same f32 rounding rule."""
    import scipy.sparse as sp

    m = sp.csr_matrix((csr.data, csr.indices, csr.indptr), shape=csr.shape)
    data = m.data.astype(np.float32).copy()
    if target_sum:
        rs = np.asarray(m.sum(axis=1, dtype=np.float64)).ravel()
        f = np.where(rs > 0, target_sum / np.where(rs > 0, rs, 1.0), 1.0)
        rows = np.repeat(np.arange(m.shape[0]), np.diff(m.indptr))
        data = (data.astype(np.float64) * f[rows]).astype(np.float32)
    if log1p:
        data = np.log1p(data)
    return sp.csr_matrix((data.astype(np.float64), m.indices, m.indptr), shape=m.shape).toarray()


def dense_pca(xd: np.ndarray, n_comps: int, zero_center: bool = True):
    """Exact PCA by LAPACK SVD (the dense FP64 oracle)."""
    xc = xd - xd.mean(axis=0) if zero_center else xd
    u, s, vt = np.linalg.svd(xc, full_matrices=False)
    n = xd.shape[0]
    return {
        "sigma": s[:n_comps],
        "embeddings": u[:, :n_comps] * s[:n_comps],
        "components": vt[:n_comps],
        "variance": s[:n_comps] ** 2 / max(n - 1, 1),
        "total_variance": float((xc ** 2).sum() / max(n - 1, 1)),
    }
