/*
 * CPU oracle for the randomized sparse PCA path — TEST INFRASTRUCTURE ONLY.
 *
 * A plain-C restatement of scx-accel's CPU randomized PCA
 * (scx-accel/src/pca/cpu.rs) and of the inputs it consumes:
 *   - the Ω stream: rand_chacha 0.3.1 ChaCha8Rng::seed_from_u64 (rand_core 0.6.4
 *     PCG32 seed expansion), rand_core BlockRng next_u32/next_u64, and
 *     rand_distr 0.4.3 StandardNormal (256-box ziggurat), as called by
 *     random_gaussian (cpu.rs:1174-1178);
 *   - the lazy normalize_total -> log1p transform
 *     (pyscx/src/lazy_transform/transforms.rs:15-72);
 *   - the f64 column statistics (scx-sparse/src/csr.rs:543-590), forward and
 *     transpose SpMM (cpu.rs:1185-1523), Householder thin QR, thin SVD,
 *     build_pca_result (cpu.rs:1530-1585) and the guarded total variance
 *     (cpu.rs:1029-1171, scx-sparse/src/moments.rs:288-306).
 *
 * Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may load
 * this library, and only as the checker / the timed CPU baseline. The product
 * path (scx_amd/) never links or calls it.
 *
 * Parity status of each part is stated beside its declaration.
 */
#ifndef SCXPCA_ORACLE_H
#define SCXPCA_ORACLE_H

#include <stdint.h>
#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- RNG: pinned by the reference KAT seeded_permutation(16, 42)
 *      (scx-ops/src/shuffle_order_tests.rs:67-80) for ChaCha8 + seed_from_u64
 *      + next_u32 + gen_range(u32) + SliceRandom::shuffle. */
typedef struct {
    uint32_t key[8];
    uint64_t counter;   /* block counter (words 12-13) */
    uint64_t stream;    /* nonce (words 14-15), zero for ChaCha8Rng::from_seed */
    uint32_t buf[64];   /* four 16-word blocks, as rand_chacha refills them */
    uint32_t index;     /* BlockRng read position, 64 = empty */
} orc_chacha8;

/* Raw ChaCha8 block (original 64-bit counter / 64-bit nonce layout). */
void     orc_chacha8_block(const uint32_t key[8], uint64_t counter, uint64_t nonce, uint32_t out[16]);
void     orc_chacha8_seed_from_u64(orc_chacha8 *rng, uint64_t seed);
uint32_t orc_chacha8_next_u32(orc_chacha8 *rng);
uint64_t orc_chacha8_next_u64(orc_chacha8 *rng);
uint64_t orc_splitmix64(uint64_t x);
void     orc_seeded_permutation(uint64_t n, uint64_t seed, uint64_t *perm_out);
/* First n raw u32 / u64 words of ChaCha8Rng::seed_from_u64(seed). */
void     orc_chacha8_u32_stream(uint64_t seed, uint64_t n, uint32_t *out);
void     orc_chacha8_u64_stream(uint64_t seed, uint64_t n, uint64_t *out);

/* ---- StandardNormal: ziggurat tables regenerated from Doornik (2005) with
 *      R = 3.6541528853610088, V = 0.00492867323399, rounded through "%.18f"
 *      as the upstream table generator printed them. Parity: partially pinned
 *      (table invariants, distribution tests); the bit-level sample stream has
 *      no reference fixture. */
void   orc_zig_norm_tables(double *x_out /*257*/, double *f_out /*257*/);
double orc_standard_normal(orc_chacha8 *rng);
/* random_gaussian(rows, cols, seed): row-major rows x cols, draw i -> element i. */
void   orc_random_gaussian(uint64_t rows, uint64_t cols, uint64_t seed, double *out);

/* ---- CSR (i64 indptr, u32 indices, f32 data). */
typedef struct {
    int64_t n_obs, n_vars;
    const int64_t  *indptr;
    const uint32_t *indices;
    const float    *data;
} orc_csr;

/* Row sums: sequential f64 left fold (scx-sparse/src/csr.rs:543-552). */
void orc_row_sums(const orc_csr *x, double *out);
/* X~ values (transforms.rs:21-72): normalize (target_sum>0) and/or log1p.
 * target_sum <= 0 disables normalize_total. */
void orc_transform(const orc_csr *x, double target_sum, int log1p, float *data_out);
/* Column sums and sums of squares in nnz order (csr.rs:570-580). */
void orc_col_sums_and_sum_sq(const orc_csr *x, double *sums, double *sum_sq);

/* Forward SpMM Y = (X - 1 mu^T) M, M row-major n_vars x k, Y row-major n_obs x k
 * (cpu.rs:1440-1523). means may be NULL. nthreads <= 0 -> OpenMP default. */
void orc_spmm_forward(const orc_csr *x, const double *m, int64_t k, const double *means,
                      double *y, int nthreads);
/* Transpose SpMM Z = (X - 1 mu^T)^T Q, Q row-major n_obs x k, Z row-major
 * n_vars x k (cpu.rs:1238-1431). Output-partitioned: bits independent of nthreads. */
void orc_spmm_transpose(const orc_csr *x, const double *q, int64_t k, const double *means,
                        double *z, int nthreads);

/* Householder thin QR of row-major A (m x n, m >= n): Q row-major m x n. */
void orc_qr_thin_q(const double *a, int64_t m, int64_t n, double *q, int nthreads);
/* Thin SVD (one-sided Jacobi) of row-major B (m x n, m >= n):
 * U row-major m x n, s[n] descending, vt row-major n x n. */
int  orc_thin_svd(const double *b, int64_t m, int64_t n, double *u, double *s, double *vt);

/* Guarded total variance (cpu.rs:1029-1171). x must hold X~ values.
 * means NULL -> uncentered. *recomputed set to 1 when the stable pass ran. */
double orc_total_variance(const orc_csr *x, const double *col_sum_sq, const double *means,
                          int *recomputed);

typedef struct {
    int64_t n_comps, n_oversamples, n_power_iterations;
    int32_t zero_center;
    uint64_t seed;
    double  target_sum;   /* <= 0: no normalize_total */
    int32_t log1p;
    int32_t nthreads;     /* <= 0: OpenMP default */
} orc_pca_params;

/* ---- covariance route (cpu.rs:198-291): lower triangle of X^T X (column-
 * major) and column sums, f64; 1 if a row is not strictly increasing. */
int orc_covariance_gram(const orc_csr *x, double *cov, double *col_sums, int nthreads);

/* Full randomized_pca_inmemory (cpu.rs:887-952) on the raw CSR with the lazy
 * transform applied once. Outputs (caller-allocated):
 *   embeddings n_obs x n_comps (row-major f64), components n_comps x n_vars,
 *   variance[n_comps], variance_ratio[n_comps], mean[n_vars] (ignored when
 *   !zero_center), sigma_all[k] (all k singular values of B).
 * Returns 0 on success, 1..3 for the validate_inputs errors (cpu.rs:979-997),
 * 4 on an SVD failure. */
int orc_randomized_pca(const orc_csr *x, const orc_pca_params *p, double *embeddings,
                       double *components, double *variance, double *variance_ratio,
                       double *mean, double *sigma_all);

/* ---- PFlog v4 (scx-accel/src/pflog.rs, pca/cpu.rs:611-866).
 * delta = log1p(4 alpha x) on x's pattern. form 0: the lazy Scale{4a} -> Log1p
 * chain, ((x as f64 * 4a) as f32).ln_1p() (transforms.rs:109-128); form 1: the
 * in-memory builder, (4a * x as f64).ln_1p() as f32 (pyscx/src/accel/pflog.rs:375-397). */
void orc_pflog_delta(const orc_csr *x, double four_alpha, int form, float *out);
/* b_i = -(sum_j delta_ij) * (1 / n_vars), f64 left fold per row
 * (pflog_baseline_from_delta, pflog.rs:186-222; csr.rs:543-552). */
void orc_pflog_baseline(const orc_csr *delta, double *out);
/* pflog_total_variance (cpu.rs:702-769); colmean_delta NULL: uncentered. */
double orc_pflog_total_variance(const double *col_sum_sq, const double *colmean_delta, const double *baseline,
                                int64_t n, int64_t nv);
/* pflog_pca_with_depth (cpu.rs:815-866) on the delta CSR (values already
 * transformed) and the baseline; outputs as orc_randomized_pca, plus the
 * total variance. target_sum / log1p of p are ignored. */
int orc_pflog_pca(const orc_csr *delta, const double *baseline, const orc_pca_params *p, double *embeddings,
                  double *components, double *variance, double *variance_ratio, double *mean,
                  double *sigma_all, double *total_variance);

#ifdef __cplusplus
}
#endif
#endif
