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


def per_pc_cosine(ref: np.ndarray, x: np.ndarray) -> np.ndarray:
    """|cos| between matching columns (sign-agnostic), the reference's GPU
    gate metric (pyscx/tests/test_accel_pca_gpu.py:176-181)."""
    a = np.asarray(ref, dtype=np.float64)
    b = np.asarray(x, dtype=np.float64)
    num = np.abs(np.sum(a * b, axis=0))
    return num / np.maximum(np.linalg.norm(a, axis=0) * np.linalg.norm(b, axis=0), 1e-300)


def exact_pca(x, kmax: int, zero_center: bool = True, device: str = "cpu"):
    """Exact top-``kmax`` PCA of a dense f64 matrix (n x nv, numpy) by the
    symmetric eigendecomposition of its Gram matrix in the smaller dimension
    (LAPACK syevd through numpy on the CPU, or through torch in f64 on the
    GPU for shapes whose CPU eigendecomposition would take minutes).
    Singular values are exact to ~eps * sigma_1^2 / sigma_j relative, far
    below the 1e-6 the comparisons need. Returns sigma (kmax), components
    (kmax x nv, orthonormal rows) and embeddings (n x kmax) = Xc V."""
    n, nv = x.shape
    if device == "cpu":
        xc = x - x.mean(axis=0) if zero_center else x
        cell_space = n < nv
        g = xc @ xc.T if cell_space else xc.T @ xc
        w, v = np.linalg.eigh(g)
        w, v = w[::-1][:kmax], v[:, ::-1][:, :kmax]
        sig = np.sqrt(np.maximum(w, 0.0))
        if cell_space:
            comps = (xc.T @ v / np.maximum(sig, 1e-300)).T
        else:
            comps = v.T
        emb = xc @ comps.T
        return {"sigma": sig, "components": comps, "embeddings": emb}
    import torch

    dev = torch.device("cuda:0")
    xt = torch.from_numpy(np.ascontiguousarray(x, dtype=np.float64)).to(dev)
    if zero_center:
        xt -= xt.mean(dim=0, keepdim=True)
    cell_space = n < nv
    g = xt @ xt.t().contiguous() if cell_space else xt.t().contiguous() @ xt
    w, v = torch.linalg.eigh(g)
    del g
    w = torch.flip(w, dims=[0])[:kmax]
    v = torch.flip(v, dims=[1])[:, :kmax].contiguous()
    sig = torch.sqrt(torch.clamp(w, min=0.0))
    if cell_space:
        comps = (xt.t().contiguous() @ v / torch.clamp(sig, min=1e-300)).t().contiguous()
    else:
        comps = v.t().contiguous()
    emb = xt @ comps.t().contiguous()
    torch.cuda.synchronize()
    out = {"sigma": sig.cpu().numpy(), "components": comps.cpu().numpy(), "embeddings": emb.cpu().numpy()}
    del xt, v, comps, emb
    torch.cuda.empty_cache()
    return out


def dense_xt(csr_x, target_sum, log1p=True):
    """Dense f64 X~ of a synth Csr with the oracle's transform (the
    reference's f32 rounding, glibc log1pf)."""
    import oracle

    v = oracle.CsrView(csr_x.indptr, csr_x.indices, csr_x.data, csr_x.shape)
    xt = oracle.transform(v, target_sum, log1p)
    n, nv = csr_x.shape
    d = np.zeros((n, nv))
    rows = np.repeat(np.arange(n), np.diff(csr_x.indptr))
    d[rows, csr_x.indices.astype(np.int64)] = xt
    return d
