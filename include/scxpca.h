/*
 * scxpca — MI355X-native randomized sparse PCA (C ABI).
 *
 * The entry points below replace, for device="gpu", the reference's GPU PCA
 * dispatch:
 *   scx_accel::randomized_pca_gpu(device_id, &source, n_components,
 *       n_oversamples, n_power_iterations, zero_center, seed, QrMethod,
 *       GpuPcaTuning) -> Result<PcaResult, AccelError>
 *   (scx-accel/src/pca/gpu.rs:32-78, which calls scx_gpu::gpu_randomized_pca,
 *    scx-gpu/src/gpu_pca.rs:706-763, and the resident operator
 *    scx-gpu/src/gpu_pca_resident.rs:87-347).
 *
 * Plain pointers and sizes only; no torch or HIP types in the signatures. A
 * Rust scx-gpu would bind these through `extern "C"` blocks; Python binds them
 * through ctypes (scx_amd/_lib.py).
 *
 * Threading: every call runs on the library's own stream of the chosen device
 * and returns after the device work completes. Calls on different matrices may
 * run from different threads. Errors never abort the process: every entry point
 * returns a status and leaves a thread-local message for scxpca_last_error.
 */
#ifndef SCXPCA_H
#define SCXPCA_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum {
    SCXPCA_OK = 0,
    SCXPCA_ERR_INVALID = 1,   /* AccelError::InvalidInput (scx-accel/src/error.rs) */
    SCXPCA_ERR_LINALG = 2,    /* AccelError::LinAlg, e.g. CholeskyQR2 non-SPD     */
    SCXPCA_ERR_OOM = 3,       /* GPU memory preflight (scx-gpu/src/gpu_pca.rs:450-462) */
    SCXPCA_ERR_DEVICE = 4,    /* HIP runtime failure                              */
    SCXPCA_ERR_NO_DEVICE = 5  /* no GPU / ordinal out of range                    */
} scxpca_status;

/* QR used for the power-iteration orthonormalization
 * (QrMethod, scx-gpu/src/cusolver.rs:256-262). */
typedef enum {
    SCXPCA_QR_HOUSEHOLDER = 0, /* stable: shifted CholeskyQR3 (no failure on ill-conditioned Y) */
    SCXPCA_QR_CHOLESKY = 1     /* CholeskyQR2; non-SPD Gram is an error, as in cusolver.rs:381-385 */
} scxpca_qr_method;

typedef enum {
    SCXPCA_PRECISION_FP32 = 0,        /* X~ streamed as f32, f32 accumulate               */
    SCXPCA_PRECISION_BF16_SKETCH = 1  /* sketch passes stream X~ as bf16, final B pass f32 */
} scxpca_precision;

/* CSR matrix: int64 indptr (n_obs+1), uint32 indices (nnz, non-decreasing per
 * row on device input; host input is sorted if needed), float32 values.
 * Raw counts; the normalize/log1p transform runs once per prepare (matrix
 * creation and each refresh), and X~ is kept in HBM in the layouts the
 * products read; the caller's arrays are never modified. */
typedef struct {
    int64_t n_obs;
    int64_t n_vars;
    const int64_t *indptr;
    const uint32_t *indices;
    const float *data;
    int32_t on_device; /* 1: device pointers on the call's device (borrowed), 0: host */
} scxpca_csr;

/* Arguments of randomized_pca_gpu (scx-accel/src/pca/gpu.rs:32) plus the lazy
 * transform chain that pyscx.accel.normalize_total/log1p record
 * (pyscx/src/lazy_transform/transforms.rs:15-72). */
typedef struct {
    int64_t n_comps;
    int64_t n_oversamples;
    int64_t n_power_iterations;
    int32_t zero_center;
    int32_t qr_method;      /* scxpca_qr_method */
    uint64_t seed;          /* ChaCha8Rng::seed_from_u64 stream, as cpu.rs:1174-1178 */
    double target_sum;      /* > 0: normalize_total to this target; <= 0: none   */
    int32_t log1p;          /* apply log1p after normalization                  */
    int32_t precision;      /* scxpca_precision                                   */
} scxpca_params;

/* Caller-allocated outputs (PcaResult, scx-accel/src/pca/cpu.rs:84-110).
 * embeddings: n_obs x n_comps row-major f32; components: n_comps x n_vars
 * row-major f32; variance / variance_ratio: n_comps f64; mean: n_vars f64 or
 * NULL; sigma: k f64 (all singular values of B) or NULL. */
typedef struct {
    float *embeddings;
    float *components;
    double *variance;
    double *variance_ratio;
    double *mean;
    double *sigma;
    int32_t on_device;       /* 1: embeddings/components are device pointers */
    /* filled by the call */
    int64_t k;               /* min(n_comps + n_oversamples, n_vars, n_obs) */
    double total_variance;
    int32_t tv_recomputed;   /* 1 if the cancellation guard re-ran the stable pass */
    double seconds_device;   /* device time of the PCA (events on the library stream) */
    int32_t resident;        /* 1: the CSR was resident in HBM; 0: streamed from host memory
                                (GpuPcaResult::resident_csr, scx-gpu/src/gpu_pca.rs:76-98) */
} scxpca_result;

typedef struct scxpca_matrix scxpca_matrix;

/* One-shot: upload (or borrow) the CSR, run the PCA, free device memory.
 * Replaces randomized_pca_gpu (scx-accel/src/pca/gpu.rs:32-78). A host CSR
 * goes through the residency pre-flight of try_build_resident_csr
 * (scx-gpu/src/gpu_pca_resident.rs:87-135: resident when this build's
 * footprint fits free HBM / 1.2 and SCX_GPU_PCA_RESIDENT is not "0"); a
 * runtime failure (out of memory, HIP error) while the resident matrix is
 * built is a decline, and the call continues on the streamed operator
 * (decline_on_runtime_failure, scx-gpu/src/gpu_pca.rs:547-550,
 * scx-gpu/src/error.rs:141-162); input defects still fail the call.
 * out->resident tells which route ran. A device CSR is always resident. */
scxpca_status scxpca_randomized_pca(const scxpca_csr *x, const scxpca_params *p,
                                    scxpca_result *out, int device);

/* Streamed: the same PCA for a host CSR larger than HBM (on_device = 0; rows
 * sorted, no repeated column in a row). Replaces the StreamingPcaOperator
 * route (scx-gpu/src/linear_operator.rs:61-266, pinned staging
 * scx-gpu/src/staging.rs): every pass moves the CSR over PCIe in shards of
 * shard_rows rows through two device slots (pinned staging, an asynchronous
 * copy stream overlapping the next shard's transfer with the current shard's
 * products); HBM holds two shards plus the n_obs x KP panels. FP32 only. */
scxpca_status scxpca_randomized_pca_streamed(const scxpca_csr *x, const scxpca_params *p, int64_t shard_rows,
                                             scxpca_result *out, int device);

/* Resident matrix: keep the CSR (and its per-row transform factors) in HBM
 * across calls. Replaces try_build_resident_csr / ResidentPcaOperator
 * (scx-gpu/src/gpu_pca_resident.rs:87-283). target_sum/log1p fix X~. */
scxpca_status scxpca_matrix_create(const scxpca_csr *x, double target_sum, int32_t log1p,
                                   int device, scxpca_matrix **out);
scxpca_status scxpca_matrix_destroy(scxpca_matrix *m);
/* The borrowed device buffers of m changed in place: recompute the row factors
 * and chunk offsets and drop cached statistics (the next PCA re-runs the stats
 * pass). No reference counterpart: the reference re-decodes shards per call. */
scxpca_status scxpca_matrix_refresh(scxpca_matrix *m);
/* PCA on a resident matrix (the transform fields of p are ignored: the
 * matrix's own are used). Replaces run_resident_power_loop + host SVD
 * (gpu_pca_resident.rs:295-347, gpu_pca.rs:592-697). */
scxpca_status scxpca_matrix_pca(scxpca_matrix *m, const scxpca_params *p, scxpca_result *out);
/* Covariance route on a resident matrix: exact PCA from the n_vars x n_vars
 * sample covariance of X~ (f64 eigendecomposition), k = n_comps; the
 * oversampling, power-iteration, seed, qr_method and precision fields are
 * ignored. Replaces covariance_pca (scx-accel/src/pca/cpu.rs:1663-1854), the
 * route CPU method="auto" takes at <= 5,000 genes (:1592,:1643-1660);
 * n_vars <= 16384. sigma (if given) receives sqrt((n_obs - 1) lambda). */
scxpca_status scxpca_matrix_pca_covariance(scxpca_matrix *m, const scxpca_params *p, scxpca_result *out);

/* PFlog (v4) normalization (Booeshaghi et al.; scx-accel/src/pflog.rs): the
 * resident operator Z = delta + b 1^T with delta = log1p(4 alpha x) on x's
 * pattern and the per-cell baseline b_i = -(1/n_vars) sum_j delta_ij.
 * delta forms: LAZY = the Scale{4a} -> Log1p chain, ((x as f64 * 4a) as
 * f32).ln_1p() (pyscx/src/lazy_transform/transforms.rs:109-128, the backed
 * and lazy X of pyscx/src/accel/pflog.rs:176-276); F64 = the in-memory
 * builder (4a * x as f64).ln_1p() as f32 (pflog.rs:375-397). */
typedef enum { SCXPCA_PFLOG_DELTA_LAZY = 0, SCXPCA_PFLOG_DELTA_F64 = 1 } scxpca_pflog_delta;
typedef struct {
    double alpha;           /* > 0 (finite): pinned; 0: estimate_alpha on the raw counts */
    double mu_min;          /* AlphaOptions::mu_min (the reference default 1e-3) */
    double fallback_alpha;  /* AlphaOptions::fallback_alpha (default 0.25) */
    int32_t delta_form;     /* scxpca_pflog_delta */
} scxpca_pflog_params;
/* AlphaEstimate (pflog.rs:268-288) / the binding's AlphaMeta (pyscx/src/accel/pflog.rs:315-327). */
typedef struct {
    double alpha;
    double pseudocount;     /* 1 / (4 alpha) */
    int64_t n_genes_used;   /* -1 when pinned */
    int32_t estimated;      /* 1: estimate_alpha, 0: pinned */
    int32_t fell_back;      /* 1: the fallback alpha was used */
} scxpca_pflog_fit;
/* Resident PFlog matrix: alpha resolved as resolve_alpha
 * (pyscx/src/accel/pflog.rs:329-350; estimate_alpha, pflog.rs:340-398, on
 * the GPU column moments of the raw counts), delta and the baseline
 * (pflog_baseline_from_delta, pflog.rs:186-222) built on the device, counts
 * checked as the reference checks them (non-finite -> error; negative ->
 * error for the F64 form). scxpca_matrix_pca on it runs pflog_pca
 * (scx-accel/src/pca/cpu.rs:815-866): the randomized PCA of Z with the two
 * rank-1 baseline terms in every SpMM and the closed-form total variance
 * (:702-769); scxpca_matrix_spmm and the projection apply Z; the covariance
 * route rejects it. */
scxpca_status scxpca_matrix_create_pflog(const scxpca_csr *x, const scxpca_pflog_params *pp, int device,
                                         scxpca_matrix **out);
/* The alpha fit and (baseline != NULL) the n_obs baseline values, host f64. */
scxpca_status scxpca_matrix_pflog_fit(scxpca_matrix *m, scxpca_pflog_fit *fit, double *baseline);
/* The dense transform Z = delta + b 1^T of a PFlog matrix, n_obs x n_vars
 * row-major f32 (each row f32(b_i), delta added at the stored entries), into
 * a host (on_device = 0) or device buffer. Replaces materialize_dense
 * (pyscx/src/accel/pflog.rs:506-533). */
scxpca_status scxpca_matrix_pflog_dense(scxpca_matrix *m, float *z, int32_t on_device);

/* PFlog PCA of a host CSR larger than HBM, streamed as
 * scxpca_randomized_pca_streamed streams (pflog_pca_with_depth,
 * scx-accel/src/pca/cpu.rs:815-866, streams its delta source): alpha pinned or
 * estimated by a streamed raw pass, every shard's delta and baseline built
 * from its own rows. fit (may be NULL) receives the alpha fit, baseline (may
 * be NULL) the n_obs baseline (host f64). out == NULL skips the PCA: the fit
 * and the baseline only (pflog's store="baseline" / "dense" on a matrix that
 * does not stay resident). The transform fields of p are ignored. FP32 only. */
scxpca_status scxpca_pflog_pca_streamed(const scxpca_csr *x, const scxpca_pflog_params *pp, const scxpca_params *p,
                                        int64_t shard_rows, scxpca_result *out, scxpca_pflog_fit *fit,
                                        double *baseline, int device);

/* Stage profiler (the reference's SCX_GPU_PROFILE stage timers,
 * scx-gpu/src/profile.rs:77-208): while enabled, every stage launch on the
 * matrix stream is bracketed by HIP events. scxpca_matrix_profile returns the
 * per-stage totals since the last call (ms_total / launches indexed by
 * scxpca_stage, either may be NULL), clears them, and sets the enable flag. */
typedef enum {
    SCXPCA_STAGE_STATS = 0,
    SCXPCA_STAGE_SPMM_FWD = 1,
    SCXPCA_STAGE_SPMM_T = 2,
    SCXPCA_STAGE_SPMM_T_REDUCE = 3,
    SCXPCA_STAGE_GRAM = 4,
    SCXPCA_STAGE_APPLY = 5,
    SCXPCA_STAGE_SPMM_FWD_BF16 = 6,
    SCXPCA_STAGE_SPMM_T_BF16 = 7,
    SCXPCA_STAGE_PREPARE = 8,   /* row factors, split offsets, tile-CSC build */
    SCXPCA_STAGE_SOLVE = 9,     /* single-workgroup Cholesky / Jacobi SVD      */
    SCXPCA_STAGE_SMALL = 10,    /* f64 CholeskyQR of the n_vars x k panels     */
    SCXPCA_N_STAGES = 11
} scxpca_stage;
scxpca_status scxpca_matrix_profile(scxpca_matrix *m, int32_t enable, double *ms_total,
                                    int64_t *launches, int32_t n_stages);
/* Restrict the stage profiler to the stages whose bit (1 << scxpca_stage) is
 * set in stage_mask; 0 selects every stage (the default). Each timed stage
 * costs two event records on the stream (a few microseconds of device time
 * each), so a caller timing one stage inside a measured loop selects only it.
 * Extends the reference's profiler, which has no per-stage switch. */
scxpca_status scxpca_matrix_profile_stages(scxpca_matrix *m, uint32_t stage_mask);

/* Stage probes on a resident matrix (host buffers in and out), used by the
 * parity tests. Each mirrors one reference function:
 *  stats: col_means_and_sum_sq (scx-format-io/src/prefetch.rs:616-650) and
 *         row sums (scx-sparse/src/csr.rs:543-552): sums, sum_sq, col_nnz
 *         (any may be NULL), row_factor = target/rowsum (1 when rowsum <= 0).
 *  spmm:  transpose == 0: Y = (X~ - 1 mu^T) M, M n_vars x k, Y n_obs x k
 *         (cpu.rs:1185-1229); transpose == 1: Z = (X~ - 1 mu^T)^T Q, Q n_obs x
 *         k, Z n_vars x k (cpu.rs:1397-1423). Row-major f64 in and out, f32 on
 *         the device. means NULL: no centering. bf16 != 0 streams X~ as bf16.
 *  qr:    orthonormal basis of row-major A (m x k) by the GPU CholeskyQR
 *         (cusolver.rs:297-409); replaces gpu_qr_q / gpu_cholesky_qr2. */
scxpca_status scxpca_matrix_stats(scxpca_matrix *m, double *sums, double *sum_sq,
                                  uint32_t *col_nnz, double *row_factor);
scxpca_status scxpca_matrix_spmm(scxpca_matrix *m, int32_t transpose, int64_t k,
                                 const double *in, const double *means, int32_t bf16,
                                 double *out);
scxpca_status scxpca_qr(const double *a, int64_t m, int64_t k, int32_t qr_method, int device,
                        double *q_out);

/* Projection Z = (X~ - 1 mu^T) W and its gradient dW = (X~ - 1 mu^T)^T dZ on a
 * resident matrix, device pointers, on the caller's HIP stream (NULL: the
 * library stream). W/dW: n_vars x k f32 row-major, Z/dZ: n_obs x k f32
 * row-major, mean: n_vars f32 or NULL. No reference counterpart (projection
 * autograd is new in this build); the kernels are the PCA's SpMMs. */
scxpca_status scxpca_project_forward(scxpca_matrix *m, const float *w, int64_t k,
                                     const float *mean, float *z, void *stream);
scxpca_status scxpca_project_backward(scxpca_matrix *m, const float *dz, int64_t k,
                                      const float *mean, float *dw, void *stream);

/* scxpca_project_backward without any device allocation. No counterpart in
 * the reference: it is for callers that cannot allocate device memory (code
 * under HIP graph capture, tight training or serving loops). The same
 * kernels and results as scxpca_project_backward; every buffer beyond dz, dw
 * and mean (the transposed SpMM's tile-range partials, the 1^T dZ sums, the
 * f64 mean, the packed dZ / dW when k is not 64 or 128, the gene order when
 * the matrix has none yet) lives in one static per-device scratch array of
 * 64 MiB, held by each GPU the library runs on until the process ends. A
 * matrix and k that need more scratch are rejected (SCXPCA_ERR_INVALID, the
 * size in the message). All calls on one GPU share that array: calls that
 * overlap on one GPU race. */
scxpca_status scxpca_project_backward_noalloc(scxpca_matrix *m, const float *dz, int64_t k,
                                              const float *mean, float *dw, void *stream);

/* Ω: ChaCha8Rng::seed_from_u64(seed) + StandardNormal, row-major rows x cols
 * f64, draw i -> element i (random_gaussian, cpu.rs:1174-1178). Host only. */
scxpca_status scxpca_gaussian(uint64_t rows, uint64_t cols, uint64_t seed, double *out);

/* The same Omega produced the way the PCA produces it on `device` (keystream
 * and draw outcomes on the GPU, the chain of draw starts on the host), copied
 * to host f64 row-major; cols <= 128. Same reference stream as above
 * (cpu.rs:1174-1178); exists so the device path can be checked bit for bit. */
scxpca_status scxpca_gaussian_device(uint64_t rows, uint64_t cols, uint64_t seed, int device, double *out);

/* Device discovery (scx-accel/src/pca/gpu.rs:81-133, device.rs:96-162). */
int scxpca_device_count(void);
scxpca_status scxpca_device_info(int device, char *name, size_t name_len, int64_t *total_mem,
                                 int32_t *n_cu);

/* Message of the last failed call on this thread; returns its length. */
size_t scxpca_last_error(char *buf, size_t len);

#ifdef __cplusplus
}
#endif
#endif
