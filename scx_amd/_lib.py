"""ctypes binding of ``scx_amd/csrc/libscxpca.so`` (C ABI in ``include/scxpca.h``).

The library is built in-tree (``python -m scx_amd.build``). Loading fails
loudly when it is missing: there is no CPU fallback on the GPU path.
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "csrc", "libscxpca.so")

_i32 = ctypes.c_int32
_i64 = ctypes.c_int64
_u64 = ctypes.c_uint64
_dbl = ctypes.c_double
_vp = ctypes.c_void_p

OK, ERR_INVALID, ERR_LINALG, ERR_OOM, ERR_DEVICE, ERR_NO_DEVICE = range(6)
QR_HOUSEHOLDER, QR_CHOLESKY = 0, 1
PRECISION_FP32, PRECISION_BF16_SKETCH = 0, 1
PFLOG_DELTA_LAZY, PFLOG_DELTA_F64 = 0, 1
ALPHA_MU_MIN, ALPHA_FALLBACK = 1e-3, 0.25  # AlphaOptions::default (scx-accel/src/pflog.rs:255-262)
STAGES = ("stats", "spmm_fwd", "spmm_t", "spmm_t_reduce", "gram", "apply", "spmm_fwd_bf16", "spmm_t_bf16",
          "prepare", "solve", "small")

# Every symbol include/scxpca.h declares.
EXPORTS = (
    "scxpca_randomized_pca",
    "scxpca_randomized_pca_streamed",
    "scxpca_pflog_pca_streamed",
    "scxpca_matrix_create",
    "scxpca_matrix_create_pflog",
    "scxpca_matrix_pflog_fit",
    "scxpca_matrix_pflog_dense",
    "scxpca_matrix_destroy",
    "scxpca_matrix_refresh",
    "scxpca_matrix_pca",
    "scxpca_matrix_pca_covariance",
    "scxpca_matrix_profile",
    "scxpca_matrix_profile_stages",
    "scxpca_matrix_stats",
    "scxpca_matrix_spmm",
    "scxpca_qr",
    "scxpca_project_forward",
    "scxpca_project_backward",
    "scxpca_project_backward_noalloc",
    "scxpca_gaussian",
    "scxpca_gaussian_device",
    "scxpca_device_count",
    "scxpca_device_info",
    "scxpca_last_error",
    "scxshard_read_info",
    "scxshard_decode_csr",
    "scxshard_encode",
)


class Csr(ctypes.Structure):
    _fields_ = [("n_obs", _i64), ("n_vars", _i64), ("indptr", _vp), ("indices", _vp), ("data", _vp),
                ("on_device", _i32)]


class Params(ctypes.Structure):
    _fields_ = [("n_comps", _i64), ("n_oversamples", _i64), ("n_power_iterations", _i64),
                ("zero_center", _i32), ("qr_method", _i32), ("seed", _u64), ("target_sum", _dbl),
                ("log1p", _i32), ("precision", _i32)]


class Result(ctypes.Structure):
    _fields_ = [("embeddings", _vp), ("components", _vp), ("variance", _vp), ("variance_ratio", _vp),
                ("mean", _vp), ("sigma", _vp), ("on_device", _i32), ("k", _i64), ("total_variance", _dbl),
                ("tv_recomputed", _i32), ("seconds_device", _dbl), ("resident", _i32)]


class PflogParams(ctypes.Structure):
    _fields_ = [("alpha", _dbl), ("mu_min", _dbl), ("fallback_alpha", _dbl), ("delta_form", _i32)]


class PflogFit(ctypes.Structure):
    _fields_ = [("alpha", _dbl), ("pseudocount", _dbl), ("n_genes_used", _i64), ("estimated", _i32),
                ("fell_back", _i32)]


class ShardInfo(ctypes.Structure):
    _fields_ = [("n_major", _i64), ("n_minor", _i64), ("nnz", _i64), ("global_offset", _i64), ("codec_id", _i32),
                ("value_encoding", _i32), ("index_dtype", _i32), ("shard_format_version", _i32),
                ("shard_type", _i32), ("n_groups", _i32)]


class ShardStats(ctypes.Structure):
    _fields_ = [("host_uploaded_bytes", _u64), ("device_decoded_bytes", _u64), ("fully_device_decoded", _i32),
                ("n_shards_scx1_gpu", ctypes.c_uint32), ("n_shards_shufdelta_gpu", ctypes.c_uint32),
                ("n_shards_host_bounced", ctypes.c_uint32), ("n_shards_other_gpu", ctypes.c_uint32),
                ("n_groups", ctypes.c_uint32), ("host_seconds", _dbl), ("upload_seconds", _dbl),
                ("device_seconds", _dbl), ("total_seconds", _dbl)]


class ScxPcaError(RuntimeError):
    def __init__(self, status: int, message: str):
        super().__init__(message)
        self.status = status


class LinAlgError(ScxPcaError):
    pass


_lib = None


def lib() -> ctypes.CDLL:
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise ImportError(
            f"native library {LIB_PATH} is not built; run `python -m scx_amd.build` "
            "(the GPU path has no CPU fallback)")
    # One HIP runtime per process: PyTorch ships its own libamdhip64 /
    # libhsa-runtime64 (same soname as /opt/rocm's). Loaded first, torch's copy
    # also serves this library; loaded after ours, torch would bring a second
    # runtime and fail to see the GPU. So torch, when installed, goes first.
    try:
        import torch  # noqa: F401
    except ImportError:
        pass
    l = ctypes.CDLL(LIB_PATH)
    for name in EXPORTS:
        getattr(l, name)  # raises AttributeError when a declared symbol is missing
    for name in EXPORTS:
        getattr(l, name).restype = ctypes.c_int
    l.scxpca_last_error.restype = ctypes.c_size_t
    l.scxpca_last_error.argtypes = [ctypes.c_char_p, ctypes.c_size_t]
    l.scxpca_randomized_pca.argtypes = [ctypes.POINTER(Csr), ctypes.POINTER(Params), ctypes.POINTER(Result), ctypes.c_int]
    l.scxpca_randomized_pca_streamed.argtypes = [ctypes.POINTER(Csr), ctypes.POINTER(Params), _i64,
                                                 ctypes.POINTER(Result), ctypes.c_int]
    l.scxpca_pflog_pca_streamed.argtypes = [ctypes.POINTER(Csr), ctypes.POINTER(PflogParams), ctypes.POINTER(Params),
                                            _i64, ctypes.POINTER(Result), ctypes.POINTER(PflogFit), _vp, ctypes.c_int]
    l.scxpca_matrix_create.argtypes = [ctypes.POINTER(Csr), _dbl, _i32, ctypes.c_int, ctypes.POINTER(_vp)]
    l.scxpca_matrix_create_pflog.argtypes = [ctypes.POINTER(Csr), ctypes.POINTER(PflogParams), ctypes.c_int,
                                             ctypes.POINTER(_vp)]
    l.scxpca_matrix_pflog_fit.argtypes = [_vp, ctypes.POINTER(PflogFit), _vp]
    l.scxpca_matrix_pflog_dense.argtypes = [_vp, _vp, _i32]
    l.scxpca_matrix_destroy.argtypes = [_vp]
    l.scxpca_matrix_refresh.argtypes = [_vp]
    l.scxpca_matrix_pca.argtypes = [_vp, ctypes.POINTER(Params), ctypes.POINTER(Result)]
    l.scxpca_matrix_pca_covariance.argtypes = [_vp, ctypes.POINTER(Params), ctypes.POINTER(Result)]
    l.scxpca_matrix_stats.argtypes = [_vp, _vp, _vp, _vp, _vp]
    l.scxpca_matrix_profile.argtypes = [_vp, _i32, _vp, _vp, _i32]
    l.scxpca_matrix_profile_stages.argtypes = [_vp, ctypes.c_uint32]
    l.scxpca_matrix_spmm.argtypes = [_vp, _i32, _i64, _vp, _vp, _i32, _vp]
    l.scxpca_qr.argtypes = [_vp, _i64, _i64, _i32, ctypes.c_int, _vp]
    l.scxpca_project_forward.argtypes = [_vp, _vp, _i64, _vp, _vp, _vp]
    l.scxpca_project_backward.argtypes = [_vp, _vp, _i64, _vp, _vp, _vp]
    l.scxpca_project_backward_noalloc.argtypes = [_vp, _vp, _i64, _vp, _vp, _vp]
    l.scxpca_gaussian.argtypes = [_u64, _u64, _u64, _vp]
    l.scxpca_gaussian_device.argtypes = [_u64, _u64, _u64, ctypes.c_int, _vp]
    l.scxpca_device_count.argtypes = []
    l.scxpca_device_info.argtypes = [ctypes.c_int, ctypes.c_char_p, ctypes.c_size_t, ctypes.POINTER(_i64),
                                     ctypes.POINTER(_i32)]
    l.scxshard_read_info.argtypes = [_vp, _u64, ctypes.POINTER(ShardInfo)]
    l.scxshard_decode_csr.argtypes = [_vp, _vp, _i64, _vp, _vp, _vp, _i64, _i64, _i32, _vp,
                                      ctypes.POINTER(ShardStats)]
    l.scxshard_encode.argtypes = [_vp, _i64, _vp, _vp, _i32, _i32, _i32, _i64, _i64, _i32, _vp, _u64,
                                  ctypes.POINTER(_u64)]
    _lib = l
    return l


def last_error() -> str:
    buf = ctypes.create_string_buffer(4096)
    lib().scxpca_last_error(buf, 4096)
    return buf.value.decode("utf-8", "replace")


def check(status: int) -> None:
    if status == OK:
        return
    msg = last_error()
    if status == ERR_INVALID:
        raise ValueError(msg)
    if status == ERR_LINALG:
        raise LinAlgError(status, msg)
    if status == ERR_OOM:
        raise MemoryError(msg)
    raise ScxPcaError(status, msg)


def ptr(a) -> ctypes.c_void_p:
    """Address of a numpy array or a torch tensor (host or device)."""
    if a is None:
        return None
    if isinstance(a, np.ndarray):
        return ctypes.c_void_p(a.ctypes.data)
    return ctypes.c_void_p(a.data_ptr())


def device_count() -> int:
    return int(lib().scxpca_device_count())


def device_info(device: int = 0) -> dict:
    name = ctypes.create_string_buffer(256)
    mem = _i64(0)
    ncu = _i32(0)
    check(lib().scxpca_device_info(device, name, 256, ctypes.byref(mem), ctypes.byref(ncu)))
    return {"name": name.value.decode(), "total_memory": int(mem.value), "compute_units": int(ncu.value)}


def gaussian(rows: int, cols: int, seed: int) -> np.ndarray:
    out = np.empty((rows, cols), dtype=np.float64)
    check(lib().scxpca_gaussian(_u64(rows), _u64(cols), _u64(seed), ptr(out)))
    return out


def gaussian_device(rows: int, cols: int, seed: int, device: int = 0) -> np.ndarray:
    """Omega as the PCA generates it on the GPU (cols <= 128)."""
    out = np.empty((rows, cols), dtype=np.float64)
    check(lib().scxpca_gaussian_device(_u64(rows), _u64(cols), _u64(seed), device, ptr(out)))
    return out


class Matrix:
    """A CSR matrix resident in HBM (owned copy of host arrays, or borrowed
    device tensors), with its normalize/log1p transform fixed. With ``pflog``
    (a dict: alpha (None: estimated), delta_form, mu_min, fallback_alpha) it
    is the PFlog operator delta + b 1^T instead."""

    def __init__(self, indptr, indices, data, shape, target_sum=None, log1p=False, device=0, pflog=None):
        self._keep = (indptr, indices, data)
        on_dev = not isinstance(indptr, np.ndarray)
        if not on_dev:
            indptr = np.ascontiguousarray(indptr, dtype=np.int64)
            indices = np.ascontiguousarray(indices, dtype=np.uint32)
            data = np.ascontiguousarray(data, dtype=np.float32)
            self._keep = (indptr, indices, data)
        self.shape = (int(shape[0]), int(shape[1]))
        self.device = device
        csr = Csr(self.shape[0], self.shape[1], ptr(indptr), ptr(indices), ptr(data), int(on_dev))
        h = _vp()
        if pflog is None:
            check(lib().scxpca_matrix_create(ctypes.byref(csr), _dbl(float(target_sum or 0.0)), int(bool(log1p)),
                                             device, ctypes.byref(h)))
        else:
            a = pflog.get("alpha")
            pp = PflogParams(0.0 if a is None else float(a), float(pflog.get("mu_min", ALPHA_MU_MIN)),
                             float(pflog.get("fallback_alpha", ALPHA_FALLBACK)),
                             int(pflog.get("delta_form", PFLOG_DELTA_LAZY)))
            check(lib().scxpca_matrix_create_pflog(ctypes.byref(csr), ctypes.byref(pp), device, ctypes.byref(h)))
        self.handle = h

    def pflog_dense(self) -> np.ndarray:
        """The dense PFlog transform Z (n_obs x n_vars, f32) of a PFlog matrix."""
        z = np.empty(self.shape, dtype=np.float32)
        check(lib().scxpca_matrix_pflog_dense(self.handle, ptr(z), 0))
        return z

    def pflog_fit(self, baseline: bool = True) -> dict:
        """The alpha fit of a PFlog matrix and (``baseline``) its per-cell baseline."""
        f = PflogFit()
        b = np.empty(self.shape[0]) if baseline else None
        check(lib().scxpca_matrix_pflog_fit(self.handle, ctypes.byref(f), ptr(b)))
        out = {"alpha": float(f.alpha), "pseudocount": float(f.pseudocount),
               "alpha_source": "estimated" if f.estimated else "pinned"}
        if f.estimated:
            out["n_genes_used"] = int(f.n_genes_used)
            out["fell_back"] = bool(f.fell_back)
        if baseline:
            out["baseline"] = b
        return out

    def close(self):
        if getattr(self, "handle", None) is not None and self.handle.value:
            check(lib().scxpca_matrix_destroy(self.handle))
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def refresh(self):
        check(lib().scxpca_matrix_refresh(self.handle))

    def profile(self, enable: bool = True, stages=None) -> dict:
        """Per-stage device time (ms) and launch counts since the last call;
        then sets profiling on or off, for the named stages only (None: all)."""
        mask = 0
        for name in stages or ():
            mask |= 1 << STAGES.index(name)
        check(lib().scxpca_matrix_profile_stages(self.handle, mask))
        n = len(STAGES)
        ms = np.zeros(n)
        cnt = np.zeros(n, dtype=np.int64)
        check(lib().scxpca_matrix_profile(self.handle, int(enable), ptr(ms), ptr(cnt), n))
        return {name: (float(ms[i]), int(cnt[i])) for i, name in enumerate(STAGES)}

    def stats(self):
        n, nv = self.shape
        s = np.empty(nv)
        ss = np.empty(nv)
        cnt = np.empty(nv, dtype=np.uint32)
        rf = np.empty(n)
        check(lib().scxpca_matrix_stats(self.handle, ptr(s), ptr(ss), ptr(cnt), ptr(rf)))
        return s, ss, cnt, rf

    def spmm(self, x: np.ndarray, transpose: bool = False, means=None, bf16: bool = False) -> np.ndarray:
        x = np.ascontiguousarray(x, dtype=np.float64)
        k = x.shape[1]
        out = np.empty((self.shape[1] if transpose else self.shape[0], k))
        mu = None if means is None else np.ascontiguousarray(means, dtype=np.float64)
        check(lib().scxpca_matrix_spmm(self.handle, int(transpose), _i64(k), ptr(x), ptr(mu), int(bf16), ptr(out)))
        return out

    def pca(self, n_comps=50, n_oversamples=10, n_power_iterations=2, zero_center=True, seed=0,
            qr_method=QR_HOUSEHOLDER, precision=PRECISION_FP32, out=None):
        """Run the randomized PCA; returns a dict of host arrays (or fills
        device tensors given in ``out`` = dict(embeddings=..., components=...))."""
        return self._run(lib().scxpca_matrix_pca, n_comps, n_oversamples, n_power_iterations, zero_center, seed,
                         qr_method, precision, out, max(1, min(n_comps + n_oversamples, *self.shape)))

    def pca_covariance(self, n_comps=50, zero_center=True, out=None):
        """Exact PCA by the covariance route (k = n_comps); same result dict."""
        return self._run(lib().scxpca_matrix_pca_covariance, n_comps, 0, 0, zero_center, 0, QR_HOUSEHOLDER,
                         PRECISION_FP32, out, max(1, n_comps))

    def _run(self, fn, n_comps, n_oversamples, n_power_iterations, zero_center, seed, qr_method, precision, out, k):
        n, nv = self.shape
        prm = Params(n_comps, n_oversamples, n_power_iterations, int(bool(zero_center)), qr_method, seed, 0.0, 0,
                     precision)
        var = np.zeros(max(n_comps, 1))
        ratio = np.zeros(max(n_comps, 1))
        mean = np.zeros(max(nv, 1))
        sigma = np.zeros(max(k, 1))
        if out is None:
            emb = np.empty((n, n_comps), dtype=np.float32)
            comps = np.empty((n_comps, nv), dtype=np.float32)
            on_dev = 0
        else:
            emb, comps = out["embeddings"], out["components"]
            on_dev = 1
        res = Result(ptr(emb), ptr(comps), ptr(var), ptr(ratio), ptr(mean), ptr(sigma), on_dev, 0, 0.0, 0, 0.0, 0)
        check(fn(self.handle, ctypes.byref(prm), ctypes.byref(res)))
        return {
            "embeddings": emb,
            "components": comps,
            "variance": var[:n_comps],
            "variance_ratio": ratio[:n_comps],
            "mean": mean[:nv] if zero_center else None,
            "sigma": sigma[: res.k],
            "k": int(res.k),
            "total_variance": float(res.total_variance),
            "tv_recomputed": bool(res.tv_recomputed),
            "seconds_device": float(res.seconds_device),
        }

    def project_forward(self, w, k, mean, z, stream=None):
        check(lib().scxpca_project_forward(self.handle, ptr(w), _i64(k), ptr(mean), ptr(z), _vp(stream or 0)))

    def project_backward(self, dz, k, mean, dw, stream=None):
        check(lib().scxpca_project_backward(self.handle, ptr(dz), _i64(k), ptr(mean), ptr(dw), _vp(stream or 0)))

    def project_backward_noalloc(self, dz, k, mean, dw, stream=None):
        """project_backward with no device allocation (a fixed per-device scratch array)."""
        check(lib().scxpca_project_backward_noalloc(self.handle, ptr(dz), _i64(k), ptr(mean), ptr(dw),
                                                    _vp(stream or 0)))


def _result_dict(res, emb, comps, var, ratio, mean, sigma, n_comps, nv, zero_center) -> dict:
    return {"embeddings": emb, "components": comps, "variance": var[:n_comps], "variance_ratio": ratio[:n_comps],
            "mean": mean[:nv] if zero_center else None, "sigma": sigma[: res.k], "k": int(res.k),
            "total_variance": float(res.total_variance), "tv_recomputed": bool(res.tv_recomputed),
            "seconds_device": float(res.seconds_device), "resident": bool(res.resident)}


def randomized_pca(indptr, indices, data, shape, n_comps=50, n_oversamples=10, n_power_iterations=2,
                   zero_center=True, seed=0, qr_method=QR_HOUSEHOLDER, target_sum=None, log1p=False,
                   precision=PRECISION_FP32, device=0) -> dict:
    """One-shot randomized PCA of a host CSR (``scxpca_randomized_pca``):
    resident when the pre-flight admits it and the build succeeds, else the
    streamed operator; the result dict of ``Matrix.pca`` plus ``resident``."""
    indptr = np.ascontiguousarray(indptr, dtype=np.int64)
    indices = np.ascontiguousarray(indices, dtype=np.uint32)
    data = np.ascontiguousarray(data, dtype=np.float32)
    n, nv = int(shape[0]), int(shape[1])
    csr = Csr(n, nv, ptr(indptr), ptr(indices), ptr(data), 0)
    prm = Params(n_comps, n_oversamples, n_power_iterations, int(bool(zero_center)), qr_method, seed,
                 float(target_sum or 0.0), int(bool(log1p)), precision)
    k = max(1, min(n_comps + n_oversamples, n, nv))
    var, ratio, mean, sigma = np.zeros(max(n_comps, 1)), np.zeros(max(n_comps, 1)), np.zeros(max(nv, 1)), np.zeros(k)
    emb = np.empty((n, n_comps), dtype=np.float32)
    comps = np.empty((n_comps, nv), dtype=np.float32)
    res = Result(ptr(emb), ptr(comps), ptr(var), ptr(ratio), ptr(mean), ptr(sigma), 0, 0, 0.0, 0, 0.0, 0)
    check(lib().scxpca_randomized_pca(ctypes.byref(csr), ctypes.byref(prm), ctypes.byref(res), device))
    return _result_dict(res, emb, comps, var, ratio, mean, sigma, n_comps, nv, zero_center)


def randomized_pca_streamed(indptr, indices, data, shape, shard_rows, n_comps=50, n_oversamples=10,
                            n_power_iterations=2, zero_center=True, seed=0, qr_method=QR_HOUSEHOLDER,
                            target_sum=None, log1p=False, device=0) -> dict:
    """The randomized PCA of a host CSR streamed to the GPU in shards of
    ``shard_rows`` rows on every pass (for matrices larger than HBM); the
    result dict of ``Matrix.pca``."""
    indptr = np.ascontiguousarray(indptr, dtype=np.int64)
    indices = np.ascontiguousarray(indices, dtype=np.uint32)
    data = np.ascontiguousarray(data, dtype=np.float32)
    n, nv = int(shape[0]), int(shape[1])
    csr = Csr(n, nv, ptr(indptr), ptr(indices), ptr(data), 0)
    prm = Params(n_comps, n_oversamples, n_power_iterations, int(bool(zero_center)), qr_method, seed,
                 float(target_sum or 0.0), int(bool(log1p)), PRECISION_FP32)
    k = max(1, min(n_comps + n_oversamples, n, nv))
    var = np.zeros(max(n_comps, 1))
    ratio = np.zeros(max(n_comps, 1))
    mean = np.zeros(max(nv, 1))
    sigma = np.zeros(k)
    emb = np.empty((n, n_comps), dtype=np.float32)
    comps = np.empty((n_comps, nv), dtype=np.float32)
    res = Result(ptr(emb), ptr(comps), ptr(var), ptr(ratio), ptr(mean), ptr(sigma), 0, 0, 0.0, 0, 0.0, 0)
    check(lib().scxpca_randomized_pca_streamed(ctypes.byref(csr), ctypes.byref(prm), _i64(int(shard_rows)),
                                               ctypes.byref(res), device))
    return {"embeddings": emb, "components": comps, "variance": var[:n_comps], "variance_ratio": ratio[:n_comps],
            "mean": mean[:nv] if zero_center else None, "sigma": sigma[: res.k], "k": int(res.k),
            "total_variance": float(res.total_variance), "tv_recomputed": bool(res.tv_recomputed),
            "seconds_device": float(res.seconds_device)}


def pflog_pca_streamed(indptr, indices, data, shape, shard_rows, alpha=None, delta_form=PFLOG_DELTA_LAZY,
                       n_comps=50, n_oversamples=10, n_power_iterations=2, zero_center=True, seed=0,
                       qr_method=QR_HOUSEHOLDER, mu_min=ALPHA_MU_MIN, fallback_alpha=ALPHA_FALLBACK, device=0,
                       pca=True):
    """PFlog PCA of a host CSR streamed to the GPU in shards on every pass;
    returns (the result dict of ``Matrix.pca``, the fit dict of
    ``Matrix.pflog_fit`` with the baseline). ``pca=False``: the fit and the
    baseline only (one streamed pass), returned as (None, fit)."""
    indptr = np.ascontiguousarray(indptr, dtype=np.int64)
    indices = np.ascontiguousarray(indices, dtype=np.uint32)
    data = np.ascontiguousarray(data, dtype=np.float32)
    n, nv = int(shape[0]), int(shape[1])
    csr = Csr(n, nv, ptr(indptr), ptr(indices), ptr(data), 0)
    pp = PflogParams(0.0 if alpha is None else float(alpha), float(mu_min), float(fallback_alpha), int(delta_form))
    prm = Params(n_comps, n_oversamples, n_power_iterations, int(bool(zero_center)), qr_method, seed, 0.0, 0,
                 PRECISION_FP32)
    base = np.zeros(n)
    f = PflogFit()
    result, res = None, None
    if pca:
        k = max(1, min(n_comps + n_oversamples, n, nv))
        var = np.zeros(max(n_comps, 1))
        ratio = np.zeros(max(n_comps, 1))
        mean = np.zeros(max(nv, 1))
        sigma = np.zeros(k)
        emb = np.empty((n, n_comps), dtype=np.float32)
        comps = np.empty((n_comps, nv), dtype=np.float32)
        res = Result(ptr(emb), ptr(comps), ptr(var), ptr(ratio), ptr(mean), ptr(sigma), 0, 0, 0.0, 0, 0.0, 0)
    check(lib().scxpca_pflog_pca_streamed(ctypes.byref(csr), ctypes.byref(pp), ctypes.byref(prm),
                                          _i64(int(shard_rows)), ctypes.byref(res) if pca else None,
                                          ctypes.byref(f), ptr(base), device))
    if pca:
        result = {"embeddings": emb, "components": comps, "variance": var[:n_comps],
                  "variance_ratio": ratio[:n_comps], "mean": mean[:nv] if zero_center else None,
                  "sigma": sigma[: res.k], "k": int(res.k), "total_variance": float(res.total_variance),
                  "tv_recomputed": bool(res.tv_recomputed), "seconds_device": float(res.seconds_device)}
    fit = {"alpha": float(f.alpha), "pseudocount": float(f.pseudocount),
           "alpha_source": "estimated" if f.estimated else "pinned", "baseline": base}
    if f.estimated:
        fit["n_genes_used"] = int(f.n_genes_used)
        fit["fell_back"] = bool(f.fell_back)
    return result, fit


def qr(a: np.ndarray, qr_method: int = QR_HOUSEHOLDER, device: int = 0) -> np.ndarray:
    a = np.ascontiguousarray(a, dtype=np.float64)
    out = np.empty_like(a)
    check(lib().scxpca_qr(ptr(a), _i64(a.shape[0]), _i64(a.shape[1]), qr_method, device, ptr(out)))
    return out
