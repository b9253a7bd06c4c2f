"""Synthetic single-cell count matrices for the PCA workloads.

Counts are negative-binomial (Gamma–Poisson) with per-gene mean
``m_g ~ LogNormal(gene_mu, gene_sigma)``, per-cell library size
``L_c ~ LogNormal(8, 0.5)`` and dispersion ``alpha`` (var = mu + alpha*mu^2):

    mu_cg = c * m_g * L_c / sum_g m_g

The scale ``c`` is not part of the workload definitions, so it is calibrated
by bisection on a 512-cell pilot so that the *expected* density
``mean(1 - (1 + alpha*mu_cg)^(-1/alpha))`` hits the requested density.

Sampling is sparse: a uniform decides whether an entry is nonzero with its
exact NB probability, and nonzero entries are drawn from the zero-truncated NB
by CDF inversion. For alpha = 0.5 the zero probability is ``1/(1+alpha*mu)^2``,
so the whole draw uses only IEEE basic operations and reproduces bit for bit
on any x86-64 host.

Streams: gene means and library sizes come from
``PCG64(SeedSequence([seed, 0]))``; row block ``b`` (``ROW_BLOCK`` rows) from
``PCG64(SeedSequence([seed, 1 + b]))``. Blocks are independent, so they are
generated in parallel and any row prefix of a matrix is the matrix of the
first rows.

Output is canonical CSR: int64 indptr, uint32 indices sorted within each row,
float32 values (integer counts).
"""
from __future__ import annotations

import os
from concurrent.futures import ProcessPoolExecutor
from dataclasses import dataclass

import numpy as np

ROW_BLOCK = 4096


@dataclass(frozen=True)
class Workload:
    name: str
    n_obs: int
    n_vars: int
    seed: int
    gene_mu: float
    gene_sigma: float
    density: float
    n_comps: int
    n_oversamples: int
    n_power_iterations: int
    alpha: float = 0.5
    lib_mu: float = 8.0
    lib_sigma: float = 0.5
    target_sum: float = 1e4


CFG1 = Workload("pca_10k_x_2k_k50", 10_000, 2_000, 0, -1.0, 1.5, 0.10, 50, 10, 2)
CFG2 = Workload("pca_1m_x_2k_k50", 1_000_000, 2_000, 0, -1.0, 1.5, 0.08, 50, 10, 3)
CFG5 = Workload("pca_2m_x_33538_k100", 2_000_000, 33_538, 1, -3.0, 2.0, 0.06, 100, 20, 4)
WORKLOADS = {w.name: w for w in (CFG1, CFG2, CFG5)}


def _zero_prob(mu, alpha: float):
    if alpha == 0.5:
        t = 1.0 + 0.5 * mu
        return 1.0 / (t * t)
    return np.power(1.0 + alpha * mu, -1.0 / alpha)


def calibrate_scale(gene_means: np.ndarray, lib: np.ndarray, density: float, alpha: float) -> float:
    """Bisection (log space, 48 steps) for the scale c giving the expected density
    on the first 512 cells."""
    w = gene_means / gene_means.sum()
    pilot = lib[:512]
    base = np.outer(pilot, w)

    def dens(c):
        return float(1.0 - _zero_prob(c * base, alpha).mean())

    lo, hi = 1e-6, 1.0
    while dens(hi) < density:
        hi *= 4.0
    for _ in range(48):
        mid = np.sqrt(lo * hi)
        if dens(mid) < density:
            lo = mid
        else:
            hi = mid
    return float(np.sqrt(lo * hi))


def _truncated_nb(rng: np.random.Generator, mu: np.ndarray, alpha: float) -> np.ndarray:
    """NB(mean mu, dispersion alpha) conditioned on x > 0, by CDF inversion."""
    r = 1.0 / alpha
    pz = _zero_prob(mu, alpha)
    u = pz + (1.0 - pz) * rng.random(mu.shape[0])
    ratio = alpha * mu / (1.0 + alpha * mu)
    k = np.zeros(mu.shape[0], dtype=np.int64)
    pk = pz.copy()
    cdf = pz.copy()
    pending = np.arange(mu.shape[0])
    while pending.size:
        kk = k[pending]
        pk[pending] = pk[pending] * (kk + r) / (kk + 1.0) * ratio[pending]
        k[pending] = kk + 1
        cdf[pending] += pk[pending]
        done = (cdf[pending] >= u[pending]) | (pk[pending] <= 0.0)
        pending = pending[~done]
    return k


@dataclass
class Csr:
    indptr: np.ndarray
    indices: np.ndarray
    data: np.ndarray
    shape: tuple
    scale: float = 0.0

    @property
    def nnz(self) -> int:
        return int(self.indptr[-1])

    def to_scipy(self):
        import scipy.sparse as sp

        return sp.csr_matrix((self.data, self.indices, self.indptr), shape=self.shape)

    def head(self, n_obs: int) -> "Csr":
        e = int(self.indptr[n_obs])
        return Csr(self.indptr[: n_obs + 1].copy(), self.indices[:e], self.data[:e],
                   (n_obs, self.shape[1]), self.scale)


def _block(args):
    seed, b, lib_blk, wgt, alpha = args
    rng = np.random.Generator(np.random.PCG64(np.random.SeedSequence([seed, 1 + b])))
    mu = np.outer(lib_blk, wgt)
    keep = rng.random(mu.shape) >= _zero_prob(mu, alpha)
    rr, cc = np.nonzero(keep)
    vals = _truncated_nb(rng, mu[rr, cc], alpha).astype(np.float32)
    counts = np.bincount(rr, minlength=lib_blk.shape[0]).astype(np.int64)
    return counts, cc.astype(np.uint32), vals


def _params(w: Workload, seed: int):
    rng = np.random.Generator(np.random.PCG64(np.random.SeedSequence([seed, 0])))
    gene_means = rng.lognormal(w.gene_mu, w.gene_sigma, size=w.n_vars)
    lib = rng.lognormal(w.lib_mu, w.lib_sigma, size=w.n_obs)
    c = calibrate_scale(gene_means, lib, w.density, w.alpha)
    return lib, (c * gene_means / gene_means.sum()), c


def generate(w: Workload, n_obs: int | None = None, seed: int | None = None, workers: int | None = None) -> Csr:
    """The workload's count matrix, or its first ``n_obs`` rows."""
    seed = w.seed if seed is None else seed
    lib, wgt, c = _params(w, seed)
    rows = w.n_obs if n_obs is None else min(n_obs, w.n_obs)
    nblk = (rows + ROW_BLOCK - 1) // ROW_BLOCK
    jobs = [(seed, b, lib[b * ROW_BLOCK : min(rows, (b + 1) * ROW_BLOCK)], wgt, w.alpha) for b in range(nblk)]
    if workers is None:
        workers = min(16, os.cpu_count() or 1)
    if workers > 1 and nblk > 1:
        with ProcessPoolExecutor(max_workers=min(workers, nblk)) as ex:
            parts = list(ex.map(_block, jobs, chunksize=1))
    else:
        parts = [_block(j) for j in jobs]
    counts = np.concatenate([p[0] for p in parts])
    indptr = np.zeros(rows + 1, dtype=np.int64)
    np.cumsum(counts, out=indptr[1:])
    indices = np.concatenate([p[1] for p in parts])
    data = np.concatenate([p[2] for p in parts])
    return Csr(indptr, indices, data, (rows, w.n_vars), c)


def random_csr(n_obs: int, n_vars: int, density: float, seed: int = 0, integer: bool = True,
               sorted_rows: bool = True) -> Csr:
    """Small uniform-random CSR for tests (counts 1..20, or floats in [0, 10))."""
    rng = np.random.Generator(np.random.PCG64(seed))
    hit = rng.random((n_obs, n_vars)) < density
    rr, cc = np.nonzero(hit)
    if integer:
        vals = rng.integers(1, 21, size=rr.size).astype(np.float32)
    else:
        vals = (rng.random(rr.size) * 10).astype(np.float32)
    counts = np.bincount(rr, minlength=n_obs)
    indptr = np.zeros(n_obs + 1, dtype=np.int64)
    indptr[1:] = np.cumsum(counts)
    indices = cc.astype(np.uint32)
    if not sorted_rows:
        for r in range(n_obs):
            a, b = indptr[r], indptr[r + 1]
            perm = rng.permutation(b - a)
            indices[a:b] = indices[a:b][perm]
            vals[a:b] = vals[a:b][perm]
    return Csr(indptr, indices, vals, (n_obs, n_vars))
