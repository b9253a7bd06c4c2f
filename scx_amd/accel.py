"""``pyscx.accel``-compatible surface for the randomized sparse PCA on MI355X.

Mirrors the reference binding so ``device="gpu"`` is a drop-in:

* ``pca(adata, n_comps=50, zero_center=True, random_state=0, n_oversamples=10,
  n_power_iterations=2, device="auto", method="auto", qr_method="householder",
  prefer_format="csr", allow_tf32=False, spmm_policy="default",
  memory_budget=None, mask_var=None)`` — pyscx/src/accel/pca.rs:573-1219,
  stub pyscx/python/pyscx/accel.pyi:318-333. Same defaults, same ValueError
  texts for bad arguments, same output slots: ``obsm["X_pca"]`` (f32),
  ``varm["PCs"]`` (f32, full var axis, masked-out genes 0), ``uns["pca"]``
  (variance, variance_ratio, backend, params) and the route stamp
  ``uns["scx_accel"]["pca"]`` (scx-accel/src/route.rs:325-372).
* ``normalize_total(adata, target_sum=None)`` / ``log1p(adata)`` record the
  transform lazily (pyscx/src/accel/preprocessing.rs:95-342,
  pyscx/src/lazy_transform/transforms.rs:15-88): ``adata.X`` becomes a
  :class:`LazyCsr`; the GPU PCA applies normalize -> log1p once per prepared
  matrix (the reference's transform, bit for bit) and keeps X~ in HBM in the
  two layouts its products read (CSR order and the row-tiled CSC). The host
  CSR is never rewritten.

Only the GPU route is implemented here. ``device="cpu"`` (and ``"auto"`` with no
GPU) raises: the CPU path of the reference is not part of this package.
"""
from __future__ import annotations

import os

import numpy as np

from . import _lib

__all__ = ["pca", "pflog", "normalize_total", "log1p", "LazyCsr", "resolve_device", "resolve_method",
           "gpu_available", "gpu_info"]

BACKEND = "scx-gpu-hip"


# ---------------------------------------------------------------------------
# device grammar (scx-accel/src/device.rs:96-162)
# ---------------------------------------------------------------------------

def gpu_available() -> bool:
    try:
        return _lib.device_count() > 0
    except ImportError:
        return False


def gpu_info(device: int = 0) -> dict:
    return _lib.device_info(device)


def resolve_device(device: str):
    """Returns ("cpu", None) or ("gpu", ordinal); raises ValueError /
    RuntimeError with the reference's messages."""
    if device == "cpu":
        return ("cpu", None)
    if device == "auto":
        return ("gpu", 0) if gpu_available() else ("cpu", None)
    if device.startswith("gpu"):
        suffix = device[3:]
        if suffix == "":
            ordinal = 0
        elif suffix.startswith(":"):
            try:
                ordinal = int(suffix[1:])
                if ordinal < 0 or not suffix[1:].isdigit():
                    raise ValueError
            except ValueError:
                raise ValueError(f"device='{device}' has a non-integer suffix; expected 'gpu:N' with N a "
                                 f"non-negative integer.") from None
        else:
            raise ValueError(f"unknown device: '{device}'. Use 'auto', 'cpu', 'gpu', or 'gpu:N'.")
        if not gpu_available():
            raise RuntimeError(f"device='{device}' requested but no ROCm GPU found")
        count = _lib.device_count()
        if ordinal >= count:
            raise RuntimeError(f"device='{device}' but only {count} ROCm device{' is' if count == 1 else 's are'} "
                               f"visible")
        return ("gpu", ordinal)
    raise ValueError(f"unknown device: '{device}'. Use 'auto', 'cpu', 'gpu', or 'gpu:N'.")


# ---------------------------------------------------------------------------
# lazy normalize/log1p view
# ---------------------------------------------------------------------------

class LazyCsr:
    """An in-memory CSR with a pending transform chain, the counterpart of the
    reference's ScxLazyTransformedDataset (pyscx/src/lazy_transform/). Only the
    fused chain the PCA path consumes is modelled: NormalizeTotal(target) then
    Log1p, or either alone."""

    def __init__(self, csr, target_sum=None, log1p=False):
        import scipy.sparse as sp

        if not sp.issparse(csr):
            csr = sp.csr_matrix(np.asarray(csr))
        csr = csr.tocsr()
        self.raw = csr
        self.target_sum = target_sum
        self.log1p = bool(log1p)

    @property
    def shape(self):
        return self.raw.shape

    def materialize(self):
        """The transformed CSR (host, f32), with the reference rounding rule
        ((x as f64 * factor) as f32).ln_1p() (transforms.rs:21-39)."""
        import scipy.sparse as sp

        m = self.raw
        data = m.data.astype(np.float32, copy=True)
        if self.target_sum is not None:
            rs = np.add.reduceat(m.data.astype(np.float64), m.indptr[:-1]) if m.nnz else np.zeros(m.shape[0])
            rs = np.where(np.diff(m.indptr) > 0, rs, 0.0)
            f = np.where(rs > 0, self.target_sum / np.where(rs > 0, rs, 1.0), 1.0)
            rows = np.repeat(np.arange(m.shape[0]), np.diff(m.indptr))
            data = (data.astype(np.float64) * f[rows]).astype(np.float32)
        if self.log1p:
            data = np.log1p(data)
        return sp.csr_matrix((data, m.indices, m.indptr), shape=m.shape)


def _median_positive_total(csr) -> float:
    rs = np.asarray(csr.sum(axis=1, dtype=np.float64)).ravel()
    pos = np.sort(rs[rs > 0])
    if pos.size == 0:
        return 0.0
    n = pos.size
    return float(pos[n // 2]) if n % 2 else float(0.5 * (pos[n // 2 - 1] + pos[n // 2]))


def normalize_total(adata, target_sum=None, device="auto") -> None:
    """Record normalize_total lazily. target_sum=None -> median of positive
    per-cell totals (scanpy semantics, preprocessing.rs:89-92)."""
    resolve_device(device)
    x = adata.X
    if isinstance(x, LazyCsr):
        if x.target_sum is not None or x.log1p:
            # a second normalization over an already-transformed view is not a
            # chain the fused kernels model; materialize first
            x = LazyCsr(x.materialize())
        base = x.raw
    else:
        base = LazyCsr(x).raw
    t = _median_positive_total(base) if target_sum is None else float(target_sum)
    adata.X = LazyCsr(base, target_sum=t, log1p=False)


def log1p(adata, device="auto") -> None:
    resolve_device(device)
    x = adata.X
    if isinstance(x, LazyCsr):
        if x.log1p:
            x = LazyCsr(x.materialize())
        adata.X = LazyCsr(x.raw, target_sum=x.target_sum, log1p=True)
    else:
        adata.X = LazyCsr(x, log1p=True)


# ---------------------------------------------------------------------------
# pca
# ---------------------------------------------------------------------------

def _bool_vec(obj, n_vars):
    try:
        arr = np.ascontiguousarray(np.asarray(obj), dtype=bool)
    except Exception:
        raise ValueError("mask_var must be a boolean array (or a var-column name resolving to one)") from None
    if arr.ndim != 1:
        raise ValueError("mask_var must be 1-D and boolean (or a var-column name resolving to one)")
    return arr


def _resolve_mask_var(adata, mask_var, n_vars):
    """resolve_mask_var (pyscx/src/accel/pca.rs:470-526)."""
    if mask_var is not None:
        if isinstance(mask_var, str):
            var = getattr(adata, "var", None)
            try:
                col = var[mask_var]
            except Exception:
                raise ValueError(f"mask_var column '{mask_var}' not found in adata.var") from None
            mask, name = _bool_vec(col, n_vars), mask_var
        else:
            mask, name = _bool_vec(mask_var, n_vars), None
    else:
        var = getattr(adata, "var", None)
        has = False
        try:
            has = var is not None and "highly_variable" in var
        except Exception:
            has = False
        if not has:
            return None
        mask, name = _bool_vec(var["highly_variable"], n_vars), "highly_variable"
    if mask.shape[0] != n_vars:
        raise ValueError(f"mask_var length {mask.shape[0]} != n_vars {n_vars}")
    cols = np.nonzero(mask)[0].astype(np.int64)
    if cols.size == 0:
        raise ValueError("mask_var selects zero genes (all-false mask)")
    return cols, name


COVARIANCE_PCA_THRESHOLD = 5_000  # scx-accel/src/pca/cpu.rs:1592
RESIDENT_VRAM_HEADROOM = 1.2      # scx-gpu/src/gpu_pca_resident.rs:54


def _resident_bytes(n: int, nv: int, nnz: int, k: int, bf16: bool = False) -> int:
    """HBM a resident matrix and its PCA workspace take in this build (the CSR
    copy, X~, chunk indices, tile-CSC, split offsets and per-tile partials
    when they stay under 32 GiB, the n x KP panels and the n_vars x KP f64
    workspace)."""
    kp = 64 if k <= 64 else (k + 127) // 128 * 128
    tiles = (n + 255) // 256
    nsplit = (nv + 127) // 128
    b = 8 * (n + 1) + 8 * nnz                         # indptr, indices, values
    b += 8 * n + 4 * nnz + 2 * nnz + 5 * (nnz + 256)  # row factors, X~, chunk index, tile rows + values
    b += 4 * tiles * (nv + 1)                         # tile offsets
    split = 4 * n * (nsplit + 1) + 16 * tiles * nv
    if split <= 32 << 30:
        b += split
    b += 2 * 4 * n * kp + 2 * 4 * nv * kp + 4 * 8 * nv * kp + 4 * 256 * nv * kp + 8 * 1024 * kp * kp
    if k > 128:  # column-major copies of the wide route
        b += 4 * max(n, nv) * k + 16 * nv * k
    if bf16:
        b += 4 * nnz
    return int(b)


def _free_hbm(device_id: int) -> int:
    try:
        import torch

        if torch.cuda.is_available():
            return int(torch.cuda.mem_get_info(device_id)[0])
    except Exception:
        pass
    return int(_lib.device_info(device_id)["total_memory"])


def resident_fits(n: int, nv: int, nnz: int, k: int, device_id: int, bf16: bool = False) -> bool:
    """The residency pre-flight of try_build_resident_csr
    (scx-gpu/src/gpu_pca_resident.rs:87-135): resident when the matrix and
    its workspace fit free HBM / 1.2 and SCX_GPU_PCA_RESIDENT is not "0"
    (read per call here; the reference reads it once per process)."""
    if os.environ.get("SCX_GPU_PCA_RESIDENT") == "0":
        return False
    return _resident_bytes(n, nv, nnz, k, bf16) <= _free_hbm(device_id) / RESIDENT_VRAM_HEADROOM


def stream_shard_rows(n: int, nv: int, nnz: int, k: int, device_id: int) -> int:
    """Rows per shard of the streamed route: two shards' structures within an
    eighth of the free HBM, a multiple of 256 rows, at least 256."""
    budget = _free_hbm(device_id) / RESIDENT_VRAM_HEADROOM / 8
    per_row = 8 + 8 * k + 21 * (nnz / max(n, 1)) + 4 * ((nv + 127) // 128 + 1) + 20 * (nv + 1) / 256
    rows = int(budget / (2 * per_row)) // 256 * 256
    return max(256, min(rows, max(n, 1)))


def resolve_method(method: str, n_vars: int, device_kind: str = "gpu") -> str:
    """The solver a ``method`` request runs. On the GPU every request but an
    explicit ``"covariance"`` is the randomized route, as the reference's GPU
    dispatch resolves it (pyscx/src/accel/pca.rs:384-389: the native GPU core
    is randomized only); ``"covariance"`` keeps this build's exact GPU
    covariance route reachable on request. On the CPU router
    (resolve_cpu_pca_method, scx-accel/src/pca/cpu.rs:1643-1660) ``"auto"``
    is the covariance route at <= 5,000 genes."""
    if method not in ("auto", "covariance", "randomized"):
        raise ValueError(f"Invalid method={method!r}; expected 'auto', 'covariance', or 'randomized'")
    if method == "auto":
        if device_kind == "gpu":
            return "randomized"
        return "covariance" if n_vars <= COVARIANCE_PCA_THRESHOLD else "randomized"
    return method


def _runtime_failure(e: BaseException) -> bool:
    """A device-side failure (out of memory, HIP error) rather than an input
    defect: the resident build declines on it (GpuError::is_runtime_failure,
    scx-gpu/src/error.rs:77-102)."""
    return isinstance(e, MemoryError) or (isinstance(e, _lib.ScxPcaError) and e.status == _lib.ERR_DEVICE)


def _route_stamp(device_id, spmm_policy, allow_tf32, resident):
    """AccelExecutionInfo fields (scx-accel/src/route.rs:325-372)."""
    return {
        "route": "gpu_csr",
        "fallback_reason": "none",
        "chunk_size": None,
        "graph_replay": False,
        "csc_available": None,
        "shards_decoded": None,
        "shards_uploaded": None,
        "math_mode": "allow_tf32" if allow_tf32 else "strict_fp32",
        "spmm_policy": spmm_policy,
        "rapids_version": None,
        "cuml_version": None,
        "cupy_version": None,
        "transfer_mode": None,
        "device_id": device_id,
        "bytes_uploaded": None,
        "n_shards_shufdelta_gpu": None,
        "resident_csr": resident,
        "reduction": "deterministic_partials",
    }


def pca(adata, n_comps: int = 50, zero_center: bool = True, random_state: int = 0, n_oversamples: int = 10,
        n_power_iterations: int = 2, device: str = "auto", method: str = "auto", qr_method: str = "householder",
        prefer_format: str = "csr", allow_tf32: bool = False, spmm_policy: str = "default", memory_budget=None,
        mask_var=None, *, precision: str = "fp32") -> None:
    """PCA of ``adata.X`` on the GPU; results written into ``adata``.

    ``method``: "auto" and "randomized" run the randomized route, as the
    reference's GPU dispatch resolves every method (pyscx/src/accel/pca.rs:
    384-389); an explicit "covariance" runs this build's exact GPU covariance
    route (f64 eigendecomposition of the sample covariance, the algorithm of
    covariance_pca, scx-accel/src/pca/cpu.rs:1663-1854), which ignores
    ``n_oversamples``, ``n_power_iterations``, ``random_state``,
    ``qr_method`` and ``precision``.

    The randomized route keeps the CSR resident in HBM when the pre-flight
    admits it (gpu_pca_resident.rs:87-135); when it does not, or when
    building the resident matrix fails at run time (out of memory, a HIP
    error), the CSR is streamed from host memory shard by shard instead
    (decline_on_runtime_failure, scx-gpu/src/gpu_pca.rs:547-550); the route
    stamp's ``resident_csr`` records which ran.

    ``precision`` (keyword only, not in the reference): ``"fp32"`` (default)
    or ``"bf16"`` — the BF16-sketch form streams X~ as bf16 in the sketch
    passes and runs the final B pass and Rayleigh-Ritz in FP32/FP64.
    """
    kind, device_id = resolve_device(device)
    if method not in ("auto", "covariance", "randomized"):
        raise ValueError(f"Invalid method={method!r}; expected 'auto', 'covariance', or 'randomized'")
    if qr_method not in ("householder", "cholesky"):
        raise ValueError(f"Invalid qr_method={qr_method!r}; expected 'householder' or 'cholesky'")
    if spmm_policy not in ("default", "deterministic", "benchmark_once"):
        raise ValueError(f"Invalid spmm_policy={spmm_policy!r}; expected 'default', 'deterministic', "
                         f"or 'benchmark_once'")
    if prefer_format != "csr":
        raise ValueError(f"pyscx.accel.pca only supports prefer_format='csr' (got {prefer_format!r}). The "
                         f"covariance build and randomized SpMM paths are row-major; CSC offers no measurable "
                         f"speed-up and is not implemented.")
    if precision not in ("fp32", "bf16"):
        raise ValueError(f"Invalid precision={precision!r}; expected 'fp32' or 'bf16'")
    if kind != "gpu":
        raise RuntimeError("scx_amd implements the GPU route of pca only; device='cpu' (or 'auto' without a "
                           "visible ROCm GPU) is served by the reference CPU implementation")

    import scipy.sparse as sp

    x = adata.X
    lazy = x if isinstance(x, LazyCsr) else LazyCsr(x)
    raw = lazy.raw
    if not isinstance(raw, sp.csr_matrix):
        raw = sp.csr_matrix(raw)
    full_n_vars = raw.shape[1]
    masked = _resolve_mask_var(adata, mask_var, full_n_vars)
    if masked is not None:
        raw = raw[:, masked[0]]
    if not raw.has_sorted_indices:
        raw = raw.sorted_indices()
    n_obs, n_vars = raw.shape

    route = resolve_method(method, n_vars, kind)
    qr = _lib.QR_CHOLESKY if qr_method == "cholesky" else _lib.QR_HOUSEHOLDER
    seed = int(random_state) & ((1 << 64) - 1)
    k = max(1, min(n_comps + n_oversamples, n_obs, n_vars))
    resident = route == "covariance" or resident_fits(n_obs, n_vars, raw.nnz, k, device_id, precision == "bf16")
    ip, ix, dv = (raw.indptr.astype(np.int64, copy=False), raw.indices.astype(np.uint32, copy=False),
                  raw.data.astype(np.float32, copy=False))
    def streamed():
        # the streaming operator (gpu_pca.rs:566-590): the CSR stays on the host
        if precision == "bf16":
            raise ValueError("precision='bf16' needs the matrix resident in HBM; the streamed route is FP32")
        return _lib.randomized_pca_streamed(ip, ix, dv, raw.shape,
                                            stream_shard_rows(n_obs, n_vars, raw.nnz, k, device_id),
                                            n_comps=n_comps, n_oversamples=n_oversamples,
                                            n_power_iterations=n_power_iterations, zero_center=zero_center,
                                            seed=seed, qr_method=qr, target_sum=lazy.target_sum,
                                            log1p=lazy.log1p, device=device_id)

    try:
        m = None
        if resident:
            try:
                m = _lib.Matrix(ip, ix, dv, raw.shape, target_sum=lazy.target_sum, log1p=lazy.log1p,
                                device=device_id)
            except (MemoryError, _lib.ScxPcaError) as e:
                # a runtime failure while building the resident matrix declines to
                # the streamed operator; the randomized route only (the covariance
                # route needs the resident matrix)
                if not _runtime_failure(e) or route != "randomized":
                    raise
                resident = False
        if not resident:
            res = streamed()
        else:
            try:
                if route == "covariance":
                    res = m.pca_covariance(n_comps=n_comps, zero_center=zero_center)
                else:
                    res = m.pca(n_comps=n_comps, n_oversamples=n_oversamples, n_power_iterations=n_power_iterations,
                                zero_center=zero_center, seed=seed, qr_method=qr,
                                precision=_lib.PRECISION_BF16_SKETCH if precision == "bf16" else _lib.PRECISION_FP32)
            finally:
                m.close()
    except ValueError as e:
        # AccelError::InvalidInput surfaces as RuntimeError (pca.rs:845)
        raise RuntimeError(str(e)) from None
    except _lib.ScxPcaError as e:
        raise RuntimeError(f"GPU PCA failed: {e}") from None

    pcs = np.zeros((full_n_vars, n_comps), dtype=np.float32)
    if masked is not None:
        pcs[masked[0], :] = res["components"].T
    else:
        pcs[:, :] = res["components"].T
    adata.obsm["X_pca"] = res["embeddings"]
    adata.varm["PCs"] = pcs
    adata.uns["pca"] = {
        "variance": np.asarray(res["variance"], dtype=np.float64),
        "variance_ratio": np.asarray(res["variance_ratio"], dtype=np.float64),
        "backend": BACKEND,
        "params": {
            "zero_center": bool(zero_center),
            "use_highly_variable": masked is not None and masked[1] == "highly_variable",
            "mask_var": None if masked is None else masked[1],
            "n_comps": int(n_comps),
        },
    }
    container = dict(adata.uns.get("scx_accel", {}) or {})
    container["pca"] = _route_stamp(device_id, spmm_policy, allow_tf32, resident)
    adata.uns["scx_accel"] = container


# ---------------------------------------------------------------------------
# pflog
# ---------------------------------------------------------------------------

DENSE_MAX_ELEMS = 200_000_000  # pyscx/src/accel/pflog.rs:60


def _pflog_dense_by_shards(ip, ix, dv, shape, alpha, form, rows, device_id):
    """The dense PFlog layer Z = f32(b_i) + delta (n_obs x n_vars f32) of a
    matrix that does not stay resident: each shard of ``rows`` rows is built
    as its own PFlog matrix with the fitted alpha pinned (its baseline and
    delta depend on its own rows and alpha only) and fills its rows of Z."""
    n, nv = int(shape[0]), int(shape[1])
    z = np.empty((n, nv), dtype=np.float32)
    for a in range(0, n, rows):
        b = min(n, a + rows)
        lo, hi = int(ip[a]), int(ip[b])
        m = _lib.Matrix(ip[a:b + 1] - lo, ix[lo:hi], dv[lo:hi], (b - a, nv), device=device_id,
                        pflog={"alpha": alpha, "delta_form": form})
        try:
            m.pflog_fit(baseline=True)
            z[a:b] = m.pflog_dense()
        finally:
            m.close()
    return z


def pflog(adata, *, alpha=None, layer=None, store: str = "pca", n_components: int = 50, n_oversamples: int = 10,
          n_power_iterations: int = 2, zero_center: bool = True, random_state: int = 0,
          obsm_key: str = "X_pflog_pca", baseline_key: str = "pflog_baseline", layer_out=None, out=None,
          store_repr: str = "delta_baseline", shard_size=None, dense_max_elems: int = DENSE_MAX_ELEMS,
          device: str = "auto") -> None:
    """PFlog (v4) normalization of raw counts on the GPU: the signature,
    defaults, validation and outputs of ``pyscx.accel.pflog``
    (pyscx/src/accel/pflog.rs:71-311).

    ``alpha=None`` estimates the matrix-wide overdispersion from the counts
    (estimate_alpha, scx-accel/src/pflog.rs:340-398); the per-cell baseline
    goes to ``adata.obs[baseline_key]``, the fit to ``adata.uns["pflog"]``,
    and ``store`` in ("pca", "all") runs the baseline-aware randomized PCA
    (pflog_pca, scx-accel/src/pca/cpu.rs:815-866) into
    ``adata.obsm[obsm_key]`` with the singular values
    sqrt(variance (n - 1)) in ``adata.uns[f"{obsm_key}_singular_values"]``;
    ``store`` in ("dense", "all") writes the dense Z (f32) to
    ``adata.layers[layer_out or "pflog"]`` under the ``dense_max_elems`` guard.
    An in-memory ``X`` builds delta as ``(4a * x as f64).ln_1p() as f32``
    (pflog.rs:375-397); a ``LazyCsr`` without transforms as the lazy
    ``Scale{4a} -> Log1p`` chain. The SCX-file output (``out=``) is not built
    here: it raises NotImplementedError after the checks the reference makes
    first."""
    if alpha is not None and (not np.isfinite(alpha) or alpha <= 0):
        raise ValueError(f"pflog: alpha must be positive and finite, got {alpha}")
    kind, device_id = resolve_device(device)
    stores = {"pca": (True, False), "baseline": (False, False), "dense": (False, True), "all": (True, True)}
    if store not in stores:
        raise ValueError(f'pflog: unknown store {store!r}; expected "pca", "baseline", "dense", or "all"')
    want_pca, want_dense = stores[store]
    if store_repr not in ("delta_baseline", "dense"):
        raise ValueError(f'pflog: unknown store_repr {store_repr!r}; expected "delta_baseline" or "dense"')
    if out is not None and not want_dense:
        raise ValueError('pflog: out= streams the materialized transform to disk; pass store="dense" or '
                         'store="all" (got store that does not materialize)')
    if layer is not None:
        try:
            x = adata.layers[layer]
        except KeyError:
            raise ValueError(f"layer '{layer}' not found in adata.layers") from None
    else:
        x = adata.X
    if want_dense and out is not None:
        raise NotImplementedError("pflog: out= writes an SCX file (the storage layer), which scx_amd does not "
                                  "build; without out= the dense transform lands in adata.layers")
    if kind != "gpu":
        raise RuntimeError("scx_amd implements the GPU route of pflog only; device='cpu' (or 'auto' without a "
                           "visible ROCm GPU) is served by the reference CPU implementation")

    import scipy.sparse as sp

    if isinstance(x, LazyCsr):
        if x.target_sum is not None or x.log1p:
            raise ValueError("pflog requires raw counts, but X is a lazy-transformed dataset that already carries "
                             "transforms (e.g. normalize_total / log1p). Run pflog on the raw-count X instead.")
        raw, form = x.raw, _lib.PFLOG_DELTA_LAZY
    else:
        raw, form = x, _lib.PFLOG_DELTA_F64
    if not sp.issparse(raw):
        raw = sp.csr_matrix(np.asarray(raw))
    raw = sp.csr_matrix(raw)
    if not raw.has_sorted_indices:
        raw = raw.sorted_indices()

    def runtime(e):
        # the binding maps estimate_alpha / baseline / PCA failures to RuntimeError and
        # delta_from_raw_csr's count checks to ValueError (pflog.rs:329-350, 375-397, 460-470)
        return ValueError(str(e)) if str(e).startswith("pflog:") else RuntimeError(str(e))

    n_elems = raw.shape[0] * raw.shape[1]
    ip, ix, dv = (raw.indptr.astype(np.int64, copy=False), raw.indices.astype(np.uint32, copy=False),
                  raw.data.astype(np.float32, copy=False))
    k = max(1, min(n_components + n_oversamples, *raw.shape))
    seed = int(random_state) & ((1 << 64) - 1)
    res = dense = None
    if not resident_fits(raw.shape[0], raw.shape[1], raw.nnz, k, device_id):
        # the matrix does not stay in HBM (every store mode): the fit, the
        # baseline and the PCA over streamed shards; the dense layer shard by
        # shard under the same dense_max_elems guard
        rows = stream_shard_rows(raw.shape[0], raw.shape[1], raw.nnz, k, device_id)
        try:
            if want_pca:
                res, fit = _lib.pflog_pca_streamed(ip, ix, dv, raw.shape, rows, alpha=alpha, delta_form=form,
                                                   n_comps=n_components, n_oversamples=n_oversamples,
                                                   n_power_iterations=n_power_iterations, zero_center=zero_center,
                                                   seed=seed, qr_method=_lib.QR_HOUSEHOLDER, device=device_id)
            else:
                _, fit = _lib.pflog_pca_streamed(ip, ix, dv, raw.shape, rows, alpha=alpha, delta_form=form,
                                                 n_comps=1, n_oversamples=0, device=device_id, pca=False)
            if want_dense and n_elems <= dense_max_elems:
                dense = _pflog_dense_by_shards(ip, ix, dv, raw.shape, fit["alpha"], form, rows, device_id)
        except ValueError as e:
            raise runtime(e) from None
        resident = False
    else:
        resident = True
        try:
            m = _lib.Matrix(ip, ix, dv, raw.shape, device=device_id, pflog={"alpha": alpha, "delta_form": form})
        except ValueError as e:
            raise runtime(e) from None
        try:
            fit = m.pflog_fit(baseline=True)
            dense = m.pflog_dense() if want_dense and n_elems <= dense_max_elems else None
            if want_pca:
                try:
                    res = m.pca(n_comps=n_components, n_oversamples=n_oversamples,
                                n_power_iterations=n_power_iterations, zero_center=zero_center, seed=seed,
                                qr_method=_lib.QR_HOUSEHOLDER)
                except (ValueError, _lib.ScxPcaError) as e:
                    raise RuntimeError(str(e)) from None
        finally:
            m.close()

    uns = {"version": "v4", "alpha": fit["alpha"], "pseudocount": fit["pseudocount"],
           "alpha_source": fit["alpha_source"]}
    if fit["alpha_source"] == "estimated":
        uns["n_genes_used"] = fit["n_genes_used"]
        uns["fell_back"] = fit["fell_back"]
    adata.uns["pflog"] = uns
    adata.obs[baseline_key] = fit["baseline"]
    if res is not None:
        adata.obsm[obsm_key] = res["embeddings"]
        scale = max(raw.shape[0] - 1.0, 1.0)
        adata.uns[f"{obsm_key}_singular_values"] = np.sqrt(np.maximum(res["variance"] * scale, 0.0))
    if want_dense:
        if dense is None:  # the in-memory guard (pyscx/src/accel/pflog.rs:477-487)
            n_obs_, n_vars_ = raw.shape
            raise RuntimeError(
                f"pflog: dense materialization is {n_obs_}\u00d7{n_vars_} = {n_elems} elements, over the "
                f"dense_max_elems={dense_max_elems} guard. The exact PFlog transform is dense; pass out=<path.scx> "
                f"to stream it to disk (any size), use store=\"pca\" for an out-of-core embedding, or raise "
                f"dense_max_elems if you have the RAM.")
        adata.layers[layer_out or "pflog"] = dense
    container = dict(adata.uns.get("scx_accel", {}) or {})
    container["pflog"] = _route_stamp(device_id, "default", False, resident)
    adata.uns["scx_accel"] = container
