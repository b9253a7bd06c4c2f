/*
 * CPU oracle for the randomized sparse PCA path — TEST INFRASTRUCTURE ONLY.
 * Not part of the product: only tests, smoke() and bench.py's CPU-baseline leg
 * load it. Every function names the reference code it restates.
 *
 * Parity status (also recorded in the design notes):
 *   ChaCha8 / seed_from_u64 / next_u32 / gen_range / shuffle  -> pinned (reference KAT)
 *   next_u64 pairing, Standard/Open01 f64, ziggurat            -> partially pinned
 *       (table invariants + published leading table entries + distribution tests)
 *   X~ transform, column stats, SpMMs, QR, SVD, assembly, TV   -> pinned against
 *       closed forms, brute-force dense f64 and numpy/LAPACK in tests/.
 */
#define _GNU_SOURCE
#include "scxpca_oracle.h"

#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

/* ------------------------------------------------------------------------ */
/* ChaCha8 (rand_chacha 0.3.1, guts.rs refill_wide): 8 rounds, 64-bit block   */
/* counter in words 12-13, 64-bit nonce in words 14-15, four consecutive      */
/* blocks per refill written block after block into a 64-word buffer.         */
/* ------------------------------------------------------------------------ */

static inline uint32_t rotl32(uint32_t x, int r) { return (x << r) | (x >> (32 - r)); }

#define QR(a, b, c, d)                 \
    do {                               \
        a += b; d ^= a; d = rotl32(d, 16); \
        c += d; b ^= c; b = rotl32(b, 12); \
        a += b; d ^= a; d = rotl32(d, 8);  \
        c += d; b ^= c; b = rotl32(b, 7);  \
    } while (0)

static void chacha8_block(const uint32_t key[8], uint64_t ctr, uint64_t nonce, uint32_t out[16]) {
    uint32_t s[16] = {0x61707865u, 0x3320646eu, 0x79622d32u, 0x6b206574u,
                      key[0], key[1], key[2], key[3], key[4], key[5], key[6], key[7],
                      (uint32_t)ctr, (uint32_t)(ctr >> 32), (uint32_t)nonce, (uint32_t)(nonce >> 32)};
    uint32_t x[16];
    memcpy(x, s, sizeof x);
    for (int r = 0; r < 4; ++r) { /* 4 double rounds = ChaCha8 */
        QR(x[0], x[4], x[8], x[12]);
        QR(x[1], x[5], x[9], x[13]);
        QR(x[2], x[6], x[10], x[14]);
        QR(x[3], x[7], x[11], x[15]);
        QR(x[0], x[5], x[10], x[15]);
        QR(x[1], x[6], x[11], x[12]);
        QR(x[2], x[7], x[8], x[13]);
        QR(x[3], x[4], x[9], x[14]);
    }
    for (int i = 0; i < 16; ++i) out[i] = x[i] + s[i];
}

void orc_chacha8_block(const uint32_t key[8], uint64_t counter, uint64_t nonce, uint32_t out[16]) {
    chacha8_block(key, counter, nonce, out);
}

static void chacha8_refill(orc_chacha8 *r) {
    for (int b = 0; b < 4; ++b) chacha8_block(r->key, r->counter + (uint64_t)b, r->stream, r->buf + 16 * b);
    r->counter += 4;
}

/* rand_core 0.6.4 SeedableRng::seed_from_u64: PCG32 (MUL 6364136223846793005,
 * INC 11634580027462260723, state advanced before output, XSH-RR output) fills
 * the 32-byte seed four little-endian bytes at a time. */
void orc_chacha8_seed_from_u64(orc_chacha8 *r, uint64_t state) {
    const uint64_t MUL = 6364136223846793005ull, INC = 11634580027462260723ull;
    for (int i = 0; i < 8; ++i) {
        state = state * MUL + INC;
        uint32_t xorshifted = (uint32_t)(((state >> 18) ^ state) >> 27);
        uint32_t rot = (uint32_t)(state >> 59);
        r->key[i] = (xorshifted >> rot) | (xorshifted << ((32u - rot) & 31u));
    }
    r->counter = 0;
    r->stream = 0;
    r->index = 64;
}

/* rand_core 0.6.4 BlockRng::next_u32. */
uint32_t orc_chacha8_next_u32(orc_chacha8 *r) {
    if (r->index >= 64) { chacha8_refill(r); r->index = 0; }
    return r->buf[r->index++];
}

/* rand_core 0.6.4 BlockRng::next_u64: low word first; a read that straddles the
 * buffer end takes the last word of the old buffer as the low half. */
uint64_t orc_chacha8_next_u64(orc_chacha8 *r) {
    uint32_t idx = r->index;
    if (idx < 63) {
        r->index += 2;
        return ((uint64_t)r->buf[idx + 1] << 32) | r->buf[idx];
    } else if (idx >= 64) {
        chacha8_refill(r);
        r->index = 2;
        return ((uint64_t)r->buf[1] << 32) | r->buf[0];
    } else {
        uint64_t lo = r->buf[63];
        chacha8_refill(r);
        r->index = 1;
        return ((uint64_t)r->buf[0] << 32) | lo;
    }
}

void orc_chacha8_u32_stream(uint64_t seed, uint64_t n, uint32_t *out) {
    orc_chacha8 r;
    orc_chacha8_seed_from_u64(&r, seed);
    for (uint64_t i = 0; i < n; ++i) out[i] = orc_chacha8_next_u32(&r);
}

void orc_chacha8_u64_stream(uint64_t seed, uint64_t n, uint64_t *out) {
    orc_chacha8 r;
    orc_chacha8_seed_from_u64(&r, seed);
    for (uint64_t i = 0; i < n; ++i) out[i] = orc_chacha8_next_u64(&r);
}

/* scx-ops/src/shuffle_order.rs:52-86 (splitmix64, permutation_seed). */
uint64_t orc_splitmix64(uint64_t x) {
    uint64_t z = x + 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

/* rand 0.8.5 UniformInt<u32>::sample_single_inclusive (Lemire widening
 * multiply with the leading-zeros zone), called through gen_range(0..ubound). */
static uint32_t gen_range_u32(orc_chacha8 *r, uint32_t ubound) {
    uint32_t range = ubound; /* high - low + 1 with low = 0, high = ubound - 1 */
    if (range == 0) return orc_chacha8_next_u32(r);
    uint32_t zone = (range << __builtin_clz(range)) - 1u;
    for (;;) {
        uint32_t v = orc_chacha8_next_u32(r);
        uint64_t m = (uint64_t)v * (uint64_t)range;
        uint32_t hi = (uint32_t)(m >> 32), lo = (uint32_t)m;
        if (lo <= zone) return hi;
    }
}

/* seeded_permutation (shuffle_order.rs:76-84) = rand 0.8.5 SliceRandom::shuffle. */
void orc_seeded_permutation(uint64_t n, uint64_t seed, uint64_t *perm) {
    const uint64_t TAG = 0x5343585348554646ull; /* b"SCXSHUFF" */
    orc_chacha8 r;
    orc_chacha8_seed_from_u64(&r, orc_splitmix64(orc_splitmix64(seed) ^ TAG));
    for (uint64_t i = 0; i < n; ++i) perm[i] = i;
    for (uint64_t i = n; i-- > 1;) {
        uint64_t j = gen_range_u32(&r, (uint32_t)(i + 1));
        uint64_t t = perm[i]; perm[i] = perm[j]; perm[j] = t;
    }
}

/* ------------------------------------------------------------------------ */
/* rand_distr 0.4.3 StandardNormal (ziggurat, symmetric, 256 boxes).          */
/* ------------------------------------------------------------------------ */

#define ZIG_LEN 256
static const double ZIG_NORM_R = 3.6541528853610088; /* printed 3.654152885361008796 */
static double zig_x[ZIG_LEN + 1], zig_f[ZIG_LEN + 1];
static int zig_ready = 0;

static double norm_f(double x) { return exp(-x * x / 2.0); }
static double norm_f_inv(double y) { return sqrt(-2.0 * log(y)); }

/* Doornik (2005) table recurrence: x0 = V/f(R), x1 = R,
 * x_i = f^-1(V/x_{i-1} + f(x_{i-1})) for i = 2..255, x256 = 0, F_i = f(x_i).
 * Each entry is rounded through "%.18f" and parsed back, as the generated
 * Rust source holds the decimal text. */
static void zig_build(void) {
    const double R = 3.6541528853610088, V = 0.00492867323399;
    double x[ZIG_LEN + 1];
    memset(x, 0, sizeof x);
    x[0] = V / norm_f(R);
    x[1] = R;
    for (int i = 2; i < ZIG_LEN; ++i) {
        double last = x[i - 1];
        x[i] = norm_f_inv(V / last + norm_f(last));
    }
    char buf[64];
    for (int i = 0; i <= ZIG_LEN; ++i) {
        snprintf(buf, sizeof buf, "%.18f", x[i]);
        zig_x[i] = strtod(buf, NULL);
        snprintf(buf, sizeof buf, "%.18f", norm_f(x[i]));
        zig_f[i] = strtod(buf, NULL);
    }
    zig_ready = 1;
}

void orc_zig_norm_tables(double *x_out, double *f_out) {
    if (!zig_ready) zig_build();
    memcpy(x_out, zig_x, sizeof zig_x);
    memcpy(f_out, zig_f, sizeof zig_f);
}

/* rand 0.8.5 IntoFloat::into_float_with_exponent for u64 -> f64. */
static inline double into_float_with_exponent(uint64_t frac52, int exponent) {
    uint64_t bits = frac52 | ((uint64_t)(1023 + exponent) << 52);
    double d;
    memcpy(&d, &bits, sizeof d);
    return d;
}

/* rand 0.8.5 Open01 for f64: (u >> 12) in [1,2) minus (1 - EPSILON/2). */
static inline double open01(orc_chacha8 *r) {
    uint64_t v = orc_chacha8_next_u64(r);
    return into_float_with_exponent(v >> 12, 0) - (1.0 - 2.220446049250313e-16 / 2.0);
}

/* rand 0.8.5 Standard for f64: (u >> 11) * 2^-53. */
static inline double standard_f64(orc_chacha8 *r) {
    uint64_t v = orc_chacha8_next_u64(r);
    return (double)(v >> 11) * (1.0 / 9007199254740992.0);
}

double orc_standard_normal(orc_chacha8 *r) {
    if (!zig_ready) zig_build();
    for (;;) {
        uint64_t bits = orc_chacha8_next_u64(r);
        unsigned i = (unsigned)(bits & 0xff);
        double u = into_float_with_exponent(bits >> 12, 1) - 3.0; /* [-1, 1) */
        double x = u * zig_x[i];
        if (fabs(x) < zig_x[i + 1]) return x;
        if (i == 0) {
            /* zero_case: tail beyond R (Marsaglia) */
            double tx = 1.0, ty = 0.0;
            while (-2.0 * ty < tx * tx) {
                double x_ = open01(r);
                double y_ = open01(r);
                tx = log(x_) / ZIG_NORM_R;
                ty = log(y_);
            }
            return u < 0.0 ? tx - ZIG_NORM_R : ZIG_NORM_R - tx;
        }
        if (zig_f[i + 1] + (zig_f[i] - zig_f[i + 1]) * standard_f64(r) < norm_f(x)) return x;
    }
}

/* random_gaussian (cpu.rs:1174-1178). */
void orc_random_gaussian(uint64_t rows, uint64_t cols, uint64_t seed, double *out) {
    orc_chacha8 r;
    orc_chacha8_seed_from_u64(&r, seed);
    uint64_t n = rows * cols;
    for (uint64_t i = 0; i < n; ++i) out[i] = orc_standard_normal(&r);
}

/* ------------------------------------------------------------------------ */
/* X~ and column statistics                                                  */
/* ------------------------------------------------------------------------ */

void orc_row_sums(const orc_csr *x, double *out) {
    for (int64_t r = 0; r < x->n_obs; ++r) {
        double s = 0.0;
        for (int64_t p = x->indptr[r]; p < x->indptr[r + 1]; ++p) s += (double)x->data[p];
        out[r] = s;
    }
}

/* transforms.rs:21-72 (fused branch) and apply_single_transform for the
 * normalize-only / log1p-only chains. */
void orc_transform(const orc_csr *x, double target_sum, int do_log1p, float *out) {
    for (int64_t r = 0; r < x->n_obs; ++r) {
        int64_t a = x->indptr[r], b = x->indptr[r + 1];
        if (target_sum > 0.0) {
            double s = 0.0;
            for (int64_t p = a; p < b; ++p) s += (double)x->data[p];
            if (s > 0.0) {
                double factor = target_sum / s;
                for (int64_t p = a; p < b; ++p) {
                    float v = (float)((double)x->data[p] * factor);
                    out[p] = do_log1p ? log1pf(v) : v;
                }
                continue;
            }
        }
        for (int64_t p = a; p < b; ++p) out[p] = do_log1p ? log1pf(x->data[p]) : x->data[p];
    }
}

void orc_col_sums_and_sum_sq(const orc_csr *x, double *sums, double *sum_sq) {
    memset(sums, 0, sizeof(double) * (size_t)x->n_vars);
    memset(sum_sq, 0, sizeof(double) * (size_t)x->n_vars);
    int64_t nnz = x->indptr[x->n_obs];
    for (int64_t p = 0; p < nnz; ++p) {
        double v = (double)x->data[p];
        uint32_t c = x->indices[p];
        sums[c] += v;
        sum_sq[c] += v * v;
    }
}

/* ------------------------------------------------------------------------ */
/* SpMM                                                                      */
/* ------------------------------------------------------------------------ */

static void set_threads(int nthreads) {
#ifdef _OPENMP
    if (nthreads > 0) omp_set_num_threads(nthreads);
#else
    (void)nthreads;
#endif
}

/* cpu.rs:1440-1523: mc = mu^T M (v outer, j inner), then per row the nonzeros
 * in stored order, then the mean correction. Row-disjoint: any thread count
 * gives the same bits. */
void orc_spmm_forward(const orc_csr *x, const double *m, int64_t k, const double *means,
                      double *y, int nthreads) {
    double *mc = NULL;
    if (means) {
        mc = (double *)calloc((size_t)k, sizeof(double));
        for (int64_t v = 0; v < x->n_vars; ++v)
            for (int64_t j = 0; j < k; ++j) mc[j] += means[v] * m[v * k + j];
    }
    set_threads(nthreads);
#pragma omp parallel for schedule(dynamic, 256)
    for (int64_t r = 0; r < x->n_obs; ++r) {
        double *yr = y + r * k;
        for (int64_t j = 0; j < k; ++j) yr[j] = 0.0;
        for (int64_t p = x->indptr[r]; p < x->indptr[r + 1]; ++p) {
            const double *mr = m + (int64_t)x->indices[p] * k;
            double val = (double)x->data[p];
            for (int64_t j = 0; j < k; ++j) yr[j] += val * mr[j];
        }
        if (mc)
            for (int64_t j = 0; j < k; ++j) yr[j] -= mc[j];
    }
    free(mc);
}

static int rows_strictly_increasing(const orc_csr *x) {
    for (int64_t r = 0; r < x->n_obs; ++r)
        for (int64_t p = x->indptr[r] + 1; p < x->indptr[r + 1]; ++p)
            if (x->indices[p] <= x->indices[p - 1]) return 0;
    return 1;
}

/* Lower bound of value v in sorted idx[0..n). */
static int64_t lower_bound_u32(const uint32_t *idx, int64_t n, uint32_t v) {
    int64_t lo = 0, hi = n;
    while (lo < hi) {
        int64_t mid = lo + (hi - lo) / 2;
        if (idx[mid] < v) lo = mid + 1; else hi = mid;
    }
    return lo;
}

/* cpu.rs:1238-1389: every output element (v, j) sums rows in ascending order
 * (stored order within a row); var blocks own disjoint slices of Z, so the
 * result is independent of the block / thread count. Then Z -= mu (1^T Q). */
void orc_spmm_transpose(const orc_csr *x, const double *q, int64_t k, const double *means,
                        double *z, int nthreads) {
    const int64_t nv = x->n_vars;
    memset(z, 0, sizeof(double) * (size_t)(nv * k));
    const int sorted = rows_strictly_increasing(x);
    set_threads(nthreads);
    int nb = 1;
#ifdef _OPENMP
    nb = omp_get_max_threads() * 4;
#endif
    if (nb > nv) nb = (int)nv;
    if (nb < 1) nb = 1;
#pragma omp parallel for schedule(dynamic, 1)
    for (int b = 0; b < nb; ++b) {
        uint32_t a0 = (uint32_t)((nv * b) / nb), a1 = (uint32_t)((nv * (b + 1)) / nb);
        if (a0 >= a1) continue;
        for (int64_t r = 0; r < x->n_obs; ++r) {
            int64_t s = x->indptr[r], e = x->indptr[r + 1];
            int64_t lo = s, hi = e;
            if (sorted) {
                lo = s + lower_bound_u32(x->indices + s, e - s, a0);
                hi = s + lower_bound_u32(x->indices + s, e - s, a1);
            }
            if (lo == hi) continue;
            const double *qr = q + r * k;
            for (int64_t p = lo; p < hi; ++p) {
                uint32_t c = x->indices[p];
                if (!sorted && (c < a0 || c >= a1)) continue;
                double val = (double)x->data[p];
                double *zr = z + (int64_t)c * k;
                for (int64_t j = 0; j < k; ++j) zr[j] += val * qr[j];
            }
        }
    }
    if (means) {
        double *sumq = (double *)calloc((size_t)k, sizeof(double));
        for (int64_t j = 0; j < k; ++j) {
            double acc = 0.0;
            for (int64_t r = 0; r < x->n_obs; ++r) acc += q[r * k + j];
            sumq[j] = acc;
        }
        for (int64_t v = 0; v < nv; ++v)
            for (int64_t j = 0; j < k; ++j) z[v * k + j] -= means[v] * sumq[j];
        free(sumq);
    }
}

/* ------------------------------------------------------------------------ */
/* Dense factorizations (f64). faer's exact bits are not reproduced; results */
/* agree at subspace / singular-value level.                                 */
/* ------------------------------------------------------------------------ */

/* In-place Householder QR of column-major A (m x n). On return the reflector
 * vectors v_j (v_j[j] normalised to 1) live below the diagonal, tau[j] holds
 * 2/(v^T v) and rdiag[j] the diagonal of R; the strict upper triangle holds R. */
static void householder_qr_inplace(double *A, int64_t m, int64_t n, double *tau, double *rdiag) {
    for (int64_t j = 0; j < n; ++j) {
        double *aj = A + j * m;
        double norm2 = 0.0;
        for (int64_t i = j; i < m; ++i) norm2 += aj[i] * aj[i];
        double norm = sqrt(norm2);
        if (norm == 0.0) { tau[j] = 0.0; rdiag[j] = 0.0; continue; }
        double alpha = aj[j] > 0 ? -norm : norm;
        double v0 = aj[j] - alpha;
        /* v = (a_j - alpha e_j) / v0 so that v[j] = 1 */
        for (int64_t i = j + 1; i < m; ++i) aj[i] /= v0;
        double vtv = 1.0;
        for (int64_t i = j + 1; i < m; ++i) vtv += aj[i] * aj[i];
        tau[j] = 2.0 / vtv;
        rdiag[j] = alpha;
#pragma omp parallel for schedule(static)
        for (int64_t c = j + 1; c < n; ++c) {
            double *ac = A + c * m;
            double w = ac[j];
            for (int64_t i = j + 1; i < m; ++i) w += aj[i] * ac[i];
            w *= tau[j];
            ac[j] -= w;
            for (int64_t i = j + 1; i < m; ++i) ac[i] -= w * aj[i];
        }
    }
}

/* Thin Q (column-major m x n) from the reflectors. */
static void householder_form_q(const double *A, int64_t m, int64_t n, const double *tau, double *Q) {
    memset(Q, 0, sizeof(double) * (size_t)(m * n));
    for (int64_t j = 0; j < n; ++j) Q[j * m + j] = 1.0;
    for (int64_t j = n - 1; j >= 0; --j) {
        if (tau[j] == 0.0) continue;
        const double *v = A + j * m;
#pragma omp parallel for schedule(static)
        for (int64_t c = j; c < n; ++c) {
            double *qc = Q + c * m;
            double w = qc[j];
            for (int64_t i = j + 1; i < m; ++i) w += v[i] * qc[i];
            w *= tau[j];
            qc[j] -= w;
            for (int64_t i = j + 1; i < m; ++i) qc[i] -= w * v[i];
        }
    }
}

void orc_qr_thin_q(const double *a, int64_t m, int64_t n, double *q, int nthreads) {
    set_threads(nthreads);
    double *A = (double *)malloc(sizeof(double) * (size_t)(m * n));
    double *Q = (double *)malloc(sizeof(double) * (size_t)(m * n));
    double *tau = (double *)malloc(sizeof(double) * (size_t)n);
    double *rd = (double *)malloc(sizeof(double) * (size_t)n);
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < m; ++i)
        for (int64_t j = 0; j < n; ++j) A[j * m + i] = a[i * n + j];
    householder_qr_inplace(A, m, n, tau, rd);
    householder_form_q(A, m, n, tau, Q);
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < m; ++i)
        for (int64_t j = 0; j < n; ++j) q[i * n + j] = Q[j * m + i];
    free(A); free(Q); free(tau); free(rd);
}

/* One-sided (Hestenes) Jacobi SVD of column-major W (n x n) in place:
 * on return W's columns are U_j * s_j and V holds the right vectors. */
static int jacobi_svd_square(double *W, double *V, int64_t n) {
    for (int64_t i = 0; i < n * n; ++i) V[i] = 0.0;
    for (int64_t i = 0; i < n; ++i) V[i * n + i] = 1.0;
    const double eps = 1e-15;
    for (int sweep = 0; sweep < 80; ++sweep) {
        int rotated = 0;
        for (int64_t p = 0; p < n - 1; ++p) {
            for (int64_t qq = p + 1; qq < n; ++qq) {
                double *wp = W + p * n, *wq = W + qq * n;
                double al = 0, be = 0, ga = 0;
                for (int64_t i = 0; i < n; ++i) { al += wp[i] * wp[i]; be += wq[i] * wq[i]; ga += wp[i] * wq[i]; }
                if (ga == 0.0 || fabs(ga) <= eps * sqrt(al * be)) continue;
                rotated = 1;
                double zeta = (be - al) / (2.0 * ga);
                double t = (zeta >= 0 ? 1.0 : -1.0) / (fabs(zeta) + sqrt(1.0 + zeta * zeta));
                double c = 1.0 / sqrt(1.0 + t * t), s = c * t;
                for (int64_t i = 0; i < n; ++i) {
                    double xp = wp[i], xq = wq[i];
                    wp[i] = c * xp - s * xq;
                    wq[i] = s * xp + c * xq;
                }
                double *vp = V + p * n, *vq = V + qq * n;
                for (int64_t i = 0; i < n; ++i) {
                    double xp = vp[i], xq = vq[i];
                    vp[i] = c * xp - s * xq;
                    vq[i] = s * xp + c * xq;
                }
            }
        }
        if (!rotated) return 0;
    }
    return 1; /* did not converge */
}

/* thin_svd_decomp (cpu.rs:458-472) via Householder QR then Jacobi on R:
 * B = Q_B R, R = U_R S V^T, U = Q_B U_R. */
int orc_thin_svd(const double *b, int64_t m, int64_t n, double *u, double *s, double *vt) {
    double *A = (double *)malloc(sizeof(double) * (size_t)(m * n));
    double *Q = (double *)malloc(sizeof(double) * (size_t)(m * n));
    double *tau = (double *)malloc(sizeof(double) * (size_t)n);
    double *rd = (double *)malloc(sizeof(double) * (size_t)n);
    double *W = (double *)calloc((size_t)(n * n), sizeof(double));
    double *V = (double *)malloc(sizeof(double) * (size_t)(n * n));
    int64_t *ord = (int64_t *)malloc(sizeof(int64_t) * (size_t)n);
    double *sv = (double *)malloc(sizeof(double) * (size_t)n);
    for (int64_t i = 0; i < m; ++i)
        for (int64_t j = 0; j < n; ++j) A[j * m + i] = b[i * n + j];
    householder_qr_inplace(A, m, n, tau, rd);
    /* R (column-major n x n) */
    for (int64_t j = 0; j < n; ++j) {
        for (int64_t i = 0; i < j; ++i) W[j * n + i] = A[j * m + i];
        W[j * n + j] = rd[j];
    }
    householder_form_q(A, m, n, tau, Q);
    int rc = jacobi_svd_square(W, V, n);
    for (int64_t j = 0; j < n; ++j) {
        double acc = 0.0;
        for (int64_t i = 0; i < n; ++i) acc += W[j * n + i] * W[j * n + i];
        sv[j] = sqrt(acc);
        ord[j] = j;
    }
    /* descending order (stable insertion sort) */
    for (int64_t i = 1; i < n; ++i) {
        int64_t t = ord[i];
        int64_t j = i;
        while (j > 0 && sv[ord[j - 1]] < sv[t]) { ord[j] = ord[j - 1]; --j; }
        ord[j] = t;
    }
    for (int64_t jj = 0; jj < n; ++jj) {
        int64_t j = ord[jj];
        s[jj] = sv[j];
        double inv = sv[j] > 0 ? 1.0 / sv[j] : 0.0;
        /* U[:, jj] = Q_B (W[:, j] / s_j) */
        for (int64_t i = 0; i < m; ++i) {
            double acc = 0.0;
            for (int64_t l = 0; l < n; ++l) acc += Q[l * m + i] * W[j * n + l];
            u[i * n + jj] = acc * inv;
        }
        for (int64_t l = 0; l < n; ++l) vt[jj * n + l] = V[j * n + l];
    }
    free(A); free(Q); free(tau); free(rd); free(W); free(V); free(ord); free(sv);
    return rc;
}

/* ------------------------------------------------------------------------ */
/* Total variance (cpu.rs:1029-1171, csr.rs:876-889, moments.rs:288-306)      */
/* ------------------------------------------------------------------------ */

double orc_total_variance(const orc_csr *x, const double *col_sum_sq, const double *means,
                          int *recomputed) {
    const int64_t n = x->n_obs, nv = x->n_vars;
    const double denom = ((double)n - 1.0) > 1.0 ? ((double)n - 1.0) : 1.0;
    if (recomputed) *recomputed = 0;
    double total = 0.0;
    if (!means) {
        for (int64_t j = 0; j < nv; ++j) total += col_sum_sq[j];
        return total / denom;
    }
    for (int64_t j = 0; j < nv; ++j) total += col_sum_sq[j] - (double)n * means[j] * means[j];
    double closed = total / denom;
    double sum_sq = 0.0, msq = 0.0;
    for (int64_t j = 0; j < nv; ++j) sum_sq += col_sum_sq[j];
    if (sum_sq <= 0.0) return closed; /* degenerate */
    for (int64_t j = 0; j < nv; ++j) msq += means[j] * means[j];
    msq *= (double)n;
    if (!(sum_sq - msq <= 1e-7 * sum_sq)) return closed; /* usable */
    /* cancelled: stable centered recompute */
    if (recomputed) *recomputed = 1;
    uint64_t *col_nnz = (uint64_t *)calloc((size_t)nv, sizeof(uint64_t));
    double t = 0.0;
    int64_t nnz = x->indptr[n];
    for (int64_t p = 0; p < nnz; ++p) {
        uint32_t c = x->indices[p];
        double d = (double)x->data[p] - means[c];
        t += d * d;
        col_nnz[c] += 1;
    }
    for (int64_t c = 0; c < nv; ++c) {
        uint64_t z = (uint64_t)n > col_nnz[c] ? (uint64_t)n - col_nnz[c] : 0;
        t += (double)z * means[c] * means[c];
    }
    free(col_nnz);
    return t / denom;
}

/* ------------------------------------------------------------------------ */
/* randomized_pca_inmemory (cpu.rs:887-952) + build_pca_result (:1530-1585)  */
/* ------------------------------------------------------------------------ */

/* Forward and transpose SpMM of the PFlog operator Z = delta + b 1^T
 * (streaming_spmm_forward_offset / _transpose_offset, cpu.rs:617-687): the
 * plain (centered) product, then the row-baseline rank-1 term. */
static void spmm_forward_offset(const orc_csr *x, const double *m, int64_t k, const double *means,
                                const double *baseline, double *y, int nthreads) {
    orc_spmm_forward(x, m, k, means, y, nthreads);
    if (!baseline) return;
    double *cs = (double *)calloc((size_t)k, sizeof(double));
    for (int64_t v = 0; v < x->n_vars; ++v)
        for (int64_t j = 0; j < k; ++j) cs[j] += m[v * k + j];
    for (int64_t i = 0; i < x->n_obs; ++i)
        for (int64_t j = 0; j < k; ++j) y[i * k + j] += baseline[i] * cs[j];
    free(cs);
}

static void spmm_transpose_offset(const orc_csr *x, const double *q, int64_t k, const double *means,
                                  const double *baseline, double *z, int nthreads) {
    orc_spmm_transpose(x, q, k, means, z, nthreads);
    if (!baseline) return;
    double *bq = (double *)calloc((size_t)k, sizeof(double));
    for (int64_t i = 0; i < x->n_obs; ++i) {
        if (baseline[i] == 0.0) continue;
        for (int64_t j = 0; j < k; ++j) bq[j] += baseline[i] * q[i * k + j];
    }
    for (int64_t v = 0; v < x->n_vars; ++v)
        for (int64_t j = 0; j < k; ++j) z[v * k + j] += bq[j];
    free(bq);
}

/* The power loop (cpu.rs:569-588, :835-857) and build_pca_result
 * (:1530-1585), shared by randomized_pca_inmemory and pflog_pca: x holds the
 * streamed values (X~, or delta for PFlog), means the centering vector or
 * NULL, baseline the PFlog row offset or NULL, tv the total variance. */
static int pca_core(const orc_csr *x, const double *means, const double *baseline, double tv,
                    const orc_pca_params *p, int64_t k, double *embeddings, double *components,
                    double *variance, double *variance_ratio, double *mean_out, double *sigma_all) {
    const int64_t n = x->n_obs, nv = x->n_vars, nc = p->n_comps;
    double *omega = (double *)malloc(sizeof(double) * (size_t)(nv * k));
    orc_random_gaussian((uint64_t)nv, (uint64_t)k, p->seed, omega);
    double *y = (double *)malloc(sizeof(double) * (size_t)(n * k));
    double *q = (double *)malloc(sizeof(double) * (size_t)(n * k));
    double *b = (double *)malloc(sizeof(double) * (size_t)(nv * k));
    double *qb = (double *)malloc(sizeof(double) * (size_t)(nv * k));

    spmm_forward_offset(x, omega, k, means, baseline, y, p->nthreads);
    orc_qr_thin_q(y, n, k, q, p->nthreads);
    for (int64_t it = 0; it < p->n_power_iterations; ++it) {
        spmm_transpose_offset(x, q, k, means, baseline, b, p->nthreads);
        if (p->n_power_iterations > 2) {
            orc_qr_thin_q(b, nv, k, qb, p->nthreads);
            spmm_forward_offset(x, qb, k, means, baseline, y, p->nthreads);
        } else {
            spmm_forward_offset(x, b, k, means, baseline, y, p->nthreads);
        }
        orc_qr_thin_q(y, n, k, q, p->nthreads);
    }
    spmm_transpose_offset(x, q, k, means, baseline, b, p->nthreads);

    double *uh = (double *)malloc(sizeof(double) * (size_t)(nv * k));
    double *sg = (double *)malloc(sizeof(double) * (size_t)k);
    double *vt = (double *)malloc(sizeof(double) * (size_t)(k * k));
    int rc = orc_thin_svd(b, nv, k, uh, sg, vt);
    if (rc == 0) {
        /* embeddings = (Q V)[:, :nc] * sigma */
#pragma omp parallel for schedule(static)
        for (int64_t i = 0; i < n; ++i) {
            const double *qi = q + i * k;
            for (int64_t j = 0; j < nc; ++j) {
                double acc = 0.0;
                for (int64_t l = 0; l < k; ++l) acc += qi[l] * vt[j * k + l];
                embeddings[i * nc + j] = acc * sg[j];
            }
        }
        for (int64_t pc = 0; pc < nc; ++pc)
            for (int64_t v = 0; v < nv; ++v) components[pc * nv + v] = uh[v * k + pc];
        const double denom = ((double)n - 1.0) > 1.0 ? ((double)n - 1.0) : 1.0;
        for (int64_t j = 0; j < nc; ++j) {
            variance[j] = sg[j] * sg[j] / denom;
            variance_ratio[j] = tv > 0.0 ? variance[j] / tv : 0.0;
        }
        if (sigma_all) memcpy(sigma_all, sg, sizeof(double) * (size_t)k);
        if (means && mean_out) memcpy(mean_out, means, sizeof(double) * (size_t)nv);
    }
    free(omega); free(y); free(q); free(b); free(qb); free(uh); free(sg); free(vt);
    return rc == 0 ? 0 : 4;
}

static int validate_inputs(int64_t n, int64_t nv, int64_t nc) {
    if (nc == 0) return 1;
    if (n == 0 || nv == 0) return 2;
    if (nc > (n < nv ? n : nv)) return 3;
    return 0;
}

static int64_t pca_k(const orc_pca_params *p, int64_t n, int64_t nv) {
    int64_t k = p->n_comps + p->n_oversamples;
    if (k > nv) k = nv;
    if (k > n) k = n;
    return k;
}

int orc_randomized_pca(const orc_csr *x0, const orc_pca_params *p, double *embeddings,
                       double *components, double *variance, double *variance_ratio,
                       double *mean_out, double *sigma_all) {
    const int64_t n = x0->n_obs, nv = x0->n_vars;
    const int v = validate_inputs(n, nv, p->n_comps);
    if (v) return v;
    set_threads(p->nthreads);
    const int64_t k = pca_k(p, n, nv);
    const int64_t nnz = x0->indptr[n];

    float *xt = (float *)malloc(sizeof(float) * (size_t)(nnz > 0 ? nnz : 1));
    orc_transform(x0, p->target_sum, p->log1p, xt);
    orc_csr x = *x0;
    x.data = xt;

    double *sums = (double *)malloc(sizeof(double) * (size_t)nv);
    double *sumsq = (double *)malloc(sizeof(double) * (size_t)nv);
    orc_col_sums_and_sum_sq(&x, sums, sumsq);
    double *means = NULL;
    if (p->zero_center) {
        means = (double *)malloc(sizeof(double) * (size_t)nv);
        for (int64_t j = 0; j < nv; ++j) means[j] = sums[j] / (double)n;
    }
    const double tv = orc_total_variance(&x, sumsq, means, NULL);
    const int rc = pca_core(&x, means, NULL, tv, p, k, embeddings, components, variance, variance_ratio,
                            mean_out, sigma_all);
    free(xt); free(sums); free(sumsq); free(means);
    return rc;
}

/* ------------------------------------------------------------------------ */
/* PFlog v4 (scx-accel/src/pflog.rs; pflog_pca, cpu.rs:611-866)               */
/* ------------------------------------------------------------------------ */

void orc_pflog_delta(const orc_csr *x, double four_alpha, int form, float *out) {
    const int64_t nnz = x->indptr[x->n_obs];
    for (int64_t p = 0; p < nnz; ++p) {
        const float v = x->data[p];
        out[p] = form ? (float)log1p(four_alpha * (double)v) : log1pf((float)((double)v * four_alpha));
    }
}

void orc_pflog_baseline(const orc_csr *delta, double *out) {
    const double inv_d = 1.0 / (double)delta->n_vars;
    for (int64_t r = 0; r < delta->n_obs; ++r) {
        double s = 0.0;
        for (int64_t p = delta->indptr[r]; p < delta->indptr[r + 1]; ++p) s += (double)delta->data[p];
        out[r] = -s * inv_d;
    }
}

double orc_pflog_total_variance(const double *col_sum_sq, const double *colmean_delta, const double *baseline,
                                int64_t n, int64_t nv) {
    double s_dd = 0.0, b2 = 0.0;
    for (int64_t j = 0; j < nv; ++j) s_dd += col_sum_sq[j];
    for (int64_t i = 0; i < n; ++i) b2 += baseline[i] * baseline[i];
    const double denom = ((double)n - 1.0) > 1.0 ? ((double)n - 1.0) : 1.0;
    const double d = (double)nv;
    double tss;
    if (!colmean_delta) {
        tss = s_dd - d * b2;
    } else {
        double bsum = 0.0;
        for (int64_t i = 0; i < n; ++i) bsum += baseline[i];
        const double bmean = bsum / ((double)n > 1.0 ? (double)n : 1.0);
        double m2 = 0.0;
        for (int64_t j = 0; j < nv; ++j) {
            const double mu = colmean_delta[j] + bmean;
            m2 += mu * mu;
        }
        tss = s_dd - d * b2 - (double)n * m2;
    }
    const double tv = tss / denom;
    return tv > 0.0 ? tv : 0.0;
}

int orc_pflog_pca(const orc_csr *delta, const double *baseline, const orc_pca_params *p, double *embeddings,
                  double *components, double *variance, double *variance_ratio, double *mean_out,
                  double *sigma_all, double *total_variance) {
    const int64_t n = delta->n_obs, nv = delta->n_vars;
    const int v = validate_inputs(n, nv, p->n_comps);
    if (v) return v;
    set_threads(p->nthreads);
    const int64_t k = pca_k(p, n, nv);
    double *sums = (double *)malloc(sizeof(double) * (size_t)nv);
    double *sumsq = (double *)malloc(sizeof(double) * (size_t)nv);
    orc_col_sums_and_sum_sq(delta, sums, sumsq);
    double *cd = NULL, *means = NULL;
    if (p->zero_center) {
        cd = (double *)malloc(sizeof(double) * (size_t)nv);
        means = (double *)malloc(sizeof(double) * (size_t)nv);
        for (int64_t j = 0; j < nv; ++j) cd[j] = sums[j] / (double)n;
        double bsum = 0.0;
        for (int64_t i = 0; i < n; ++i) bsum += baseline[i];
        const double bmean = bsum / ((double)n > 1.0 ? (double)n : 1.0);
        for (int64_t j = 0; j < nv; ++j) means[j] = cd[j] + bmean;
    }
    const double tv = orc_pflog_total_variance(sumsq, cd, baseline, n, nv);
    if (total_variance) *total_variance = tv;
    const int rc = pca_core(delta, means, baseline, tv, p, k, embeddings, components, variance, variance_ratio,
                            mean_out, sigma_all);
    free(sums); free(sumsq); free(cd); free(means);
    return rc;
}


/* accumulate_covariance_block / accumulate_covariance_into (cpu.rs:198-291)
 * for a strictly increasing CSR: the lower triangle of X^T X (entry (hi, lo)
 * at cov[lo * n_vars + hi], column-major) and the column sums, f64. The
 * output is split into column blocks, one per OpenMP task; every entry adds
 * its rows in ascending order, so the result does not depend on the split or
 * the thread count (as the reference's). Returns 1 (nothing written) when a
 * row is not strictly increasing (the reference then takes its serial
 * duplicate-coalescing path, which this build's GPU path rejects). */
int orc_covariance_gram(const orc_csr *x, double *cov, double *col_sums, int nthreads) {
    const int64_t nv = x->n_vars;
    for (int64_t r = 0; r < x->n_obs; ++r)
        for (int64_t p = x->indptr[r] + 1; p < x->indptr[r + 1]; ++p)
            if (x->indices[p] <= x->indices[p - 1]) return 1;
    memset(cov, 0, sizeof(double) * (size_t)nv * (size_t)nv);
    memset(col_sums, 0, sizeof(double) * (size_t)nv);
    set_threads(nthreads);
    const int64_t nb = nv < 64 ? nv : 64;  /* column blocks */
#pragma omp parallel for schedule(dynamic, 1)
    for (int64_t b = 0; b < nb; ++b) {
        const uint32_t a0 = (uint32_t)(b * nv / nb), a1 = (uint32_t)((b + 1) * nv / nb);
        for (int64_t r = 0; r < x->n_obs; ++r) {
            const int64_t s0 = x->indptr[r], s1 = x->indptr[r + 1];
            for (int64_t i = s0; i < s1; ++i) {
                const uint32_t c1 = x->indices[i];
                if (c1 < a0) continue;
                if (c1 >= a1) break;
                const double v1 = (double)x->data[i];
                col_sums[c1] += v1;
                double *col = cov + (size_t)c1 * (size_t)nv;
                for (int64_t j = i; j < s1; ++j) col[x->indices[j]] += v1 * (double)x->data[j];
            }
        }
    }
    return 0;
}
