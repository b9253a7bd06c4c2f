// HIP kernels (gfx950) for the randomized sparse PCA.
//
// Data layout in HBM:
//   CSR          indptr i64[n_obs+1], indices u32[nnz], values f32[nnz] (raw counts)
//   rowf         f64[n_obs]   normalize_total factor target/rowsum (1 where rowsum <= 0)
//   xbf          bf16[nnz]    X~ materialized once for the BF16 sketch form
//   spl          u32[n_obs][nsplit+1]  offset of the first nonzero of each CB-gene
//                chunk inside each row (sorted rows)
//   panels       f32 row-major, KP columns (k' padded to 32/64/128, pad columns 0):
//                M (n_vars x KP), Y and Q (n_obs x KP)
//   B            f64 n_vars x KP
//
// X~ = log1p(f32(x * rowf)) is computed inside every SpMM from the raw value and
// the row factor; it is never stored in the FP32 form.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace scxpca {

constexpr int kWave = 64;
constexpr int kXfNorm = 1;
constexpr int kXfLog1p = 2;

typedef float v16f __attribute__((ext_vector_type(16)));

__device__ __forceinline__ void wave_lds_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// transforms.rs:21-39: ((x as f64 * factor) as f32).ln_1p(); the factor is 1
// for rows whose total is not positive, which reproduces the :40-72 branch.
__device__ __forceinline__ float xtilde(float v, double f, int flags) {
    const float s = (flags & kXfNorm) ? static_cast<float>(static_cast<double>(v) * f) : v;
    return (flags & kXfLog1p) ? log1pf(s) : s;
}

__device__ __forceinline__ float bf16_to_f32(uint16_t h) {
    return __uint_as_float(static_cast<uint32_t>(h) << 16);
}

__device__ __forceinline__ uint16_t f32_to_bf16_rne(float f) {
    uint32_t u = __float_as_uint(f);
    if ((u & 0x7fffffffu) > 0x7f800000u) return static_cast<uint16_t>((u >> 16) | 0x40u);  // NaN
    u += 0x7fffu + ((u >> 16) & 1u);
    return static_cast<uint16_t>(u >> 16);
}

template <typename T>
__device__ __forceinline__ T wave_sum(T v) {
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
    return v;
}

// ---------------------------------------------------------------------------
// Per-row preparation
// ---------------------------------------------------------------------------

// Row sums (f64) and the normalize_total factor; one wavefront per row.
__global__ void __launch_bounds__(256) k_row_factor(const int64_t *__restrict__ indptr,
                                                    const float *__restrict__ val, int64_t n_obs,
                                                    double target, double *__restrict__ rowf) {
    const int lane = threadIdx.x & 63;
    const int64_t wstride = static_cast<int64_t>(gridDim.x) * (blockDim.x / kWave);
    for (int64_t r = static_cast<int64_t>(blockIdx.x) * (blockDim.x / kWave) + threadIdx.x / kWave;
         r < n_obs; r += wstride) {
        const int64_t a = indptr[r], b = indptr[r + 1];
        double s = 0.0;
        for (int64_t p = a + lane; p < b; p += kWave) s += static_cast<double>(val[p]);
        s = wave_sum(s);
        if (lane == 0) rowf[r] = s > 0.0 ? target / s : 1.0;
    }
}

// Flags rows whose indices decrease (the chunked kernels need sorted rows).
__global__ void __launch_bounds__(256) k_check_sorted(const int64_t *__restrict__ indptr,
                                                      const uint32_t *__restrict__ idx, int64_t n_obs,
                                                      int64_t n_vars, int *__restrict__ flag) {
    const int64_t tid = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    const int64_t nthreads = static_cast<int64_t>(gridDim.x) * blockDim.x;
    const int64_t nnz = indptr[n_obs];
    // one thread per nonzero: compare with the previous entry of the same row
    int bad = 0;
    for (int64_t p = tid; p < nnz; p += nthreads) {
        if (idx[p] >= static_cast<uint64_t>(n_vars)) bad |= 2;
    }
    for (int64_t r = tid; r < n_obs; r += nthreads) {
        const int64_t a = indptr[r], b = indptr[r + 1];
        if (b < a) bad |= 4;
        for (int64_t p = a + 1; p < b; ++p)
            if (idx[p] < idx[p - 1]) { bad |= 1; break; }
    }
    if (bad) atomicOr(flag, bad);
}

// spl[r][g] = number of nonzeros of row r with index < g*cb, g = 0..nsplit
// (spl[r][nsplit] = row length). One wavefront per row, one lane per boundary.
__global__ void __launch_bounds__(256) k_splits(const int64_t *__restrict__ indptr,
                                                const uint32_t *__restrict__ idx, int64_t n_obs,
                                                int cb, int nsplit, uint32_t *__restrict__ spl) {
    const int lane = threadIdx.x & 63;
    const int64_t wstride = static_cast<int64_t>(gridDim.x) * (blockDim.x / kWave);
    for (int64_t r = static_cast<int64_t>(blockIdx.x) * (blockDim.x / kWave) + threadIdx.x / kWave;
         r < n_obs; r += wstride) {
        const int64_t a = indptr[r], len = indptr[r + 1] - a;
        for (int g = lane; g <= nsplit; g += kWave) {
            uint32_t pos;
            if (g == nsplit) {
                pos = static_cast<uint32_t>(len);
            } else {
                const uint32_t bound = static_cast<uint32_t>(g) * static_cast<uint32_t>(cb);
                int64_t lo = 0, hi = len;
                while (lo < hi) {
                    const int64_t mid = (lo + hi) >> 1;
                    if (idx[a + mid] < bound) lo = mid + 1; else hi = mid;
                }
                pos = static_cast<uint32_t>(lo);
            }
            spl[r * (nsplit + 1) + g] = pos;
        }
    }
}

// X~ as bf16 (RNE), materialized once for the sketch form.
__global__ void __launch_bounds__(256) k_xtilde_bf16(const int64_t *__restrict__ indptr,
                                                     const float *__restrict__ val,
                                                     const double *__restrict__ rowf, int flags,
                                                     int64_t n_obs, uint16_t *__restrict__ out) {
    const int lane = threadIdx.x & 63;
    const int64_t wstride = static_cast<int64_t>(gridDim.x) * (blockDim.x / kWave);
    for (int64_t r = static_cast<int64_t>(blockIdx.x) * (blockDim.x / kWave) + threadIdx.x / kWave;
         r < n_obs; r += wstride) {
        const int64_t a = indptr[r], b = indptr[r + 1];
        const double f = rowf ? rowf[r] : 1.0;
        for (int64_t p = a + lane; p < b; p += kWave) out[p] = f32_to_bf16_rne(xtilde(val[p], f, flags));
    }
}

// ---------------------------------------------------------------------------
// Column statistics of X~: sum, sum of squares (f64) and nonzero count per
// gene. Block (rb, gc) owns rows of row block rb and genes [gc*CG, gc*CG+CG);
// LDS accumulators, partials per row block, then a fixed-order reduction.
// ---------------------------------------------------------------------------

template <int CG>
__global__ void __launch_bounds__(256) k_col_stats(const int64_t *__restrict__ indptr,
                                                   const uint32_t *__restrict__ idx,
                                                   const float *__restrict__ val,
                                                   const double *__restrict__ rowf, int flags,
                                                   const uint32_t *__restrict__ spl, int nsplit, int cpc,
                                                   int nch, int64_t n_obs, int64_t n_vars,
                                                   int64_t rows_per_rb, double *__restrict__ psum,
                                                   double *__restrict__ psq, uint32_t *__restrict__ pcnt) {
    __shared__ double s_sum[CG];
    __shared__ double s_sq[CG];
    __shared__ uint32_t s_cnt[CG];
    const int rb = blockIdx.x / nch, gc = blockIdx.x % nch;
    const int64_t g0 = static_cast<int64_t>(gc) * CG;
    for (int i = threadIdx.x; i < CG; i += blockDim.x) { s_sum[i] = 0.0; s_sq[i] = 0.0; s_cnt[i] = 0u; }
    __syncthreads();
    const int lane = threadIdx.x & 63, wave = threadIdx.x / kWave, nw = blockDim.x / kWave;
    const int64_t r0 = static_cast<int64_t>(rb) * rows_per_rb;
    const int64_t r1 = r0 + rows_per_rb < n_obs ? r0 + rows_per_rb : n_obs;
    const int s_lo = gc * cpc, s_hi = (gc + 1) * cpc < nsplit ? (gc + 1) * cpc : nsplit;
    for (int64_t r = r0 + wave; r < r1; r += nw) {
        const int64_t a = indptr[r];
        const int64_t lo = a + spl[r * (nsplit + 1) + s_lo], hi = a + spl[r * (nsplit + 1) + s_hi];
        const double f = rowf ? rowf[r] : 1.0;
        for (int64_t p = lo + lane; p < hi; p += kWave) {
            const double x = static_cast<double>(xtilde(val[p], f, flags));
            const int c = static_cast<int>(idx[p] - g0);
            atomicAdd(&s_sum[c], x);
            atomicAdd(&s_sq[c], x * x);
            atomicAdd(&s_cnt[c], 1u);
        }
    }
    __syncthreads();
    for (int i = threadIdx.x; i < CG && g0 + i < n_vars; i += blockDim.x) {
        const int64_t o = static_cast<int64_t>(rb) * n_vars + g0 + i;
        psum[o] = s_sum[i];
        psq[o] = s_sq[i];
        pcnt[o] = s_cnt[i];
    }
}

__global__ void __launch_bounds__(256) k_col_stats_reduce(const double *__restrict__ psum,
                                                          const double *__restrict__ psq,
                                                          const uint32_t *__restrict__ pcnt, int nrb,
                                                          int64_t n_vars, double *__restrict__ sum,
                                                          double *__restrict__ sq,
                                                          uint32_t *__restrict__ cnt) {
    const int64_t v = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (v >= n_vars) return;
    double s = 0.0, q = 0.0;
    uint32_t c = 0;
    for (int rb = 0; rb < nrb; ++rb) {
        const int64_t o = static_cast<int64_t>(rb) * n_vars + v;
        s += psum[o];
        q += psq[o];
        c += pcnt[o];
    }
    sum[v] = s;
    sq[v] = q;
    cnt[v] = c;
}

// Stable centered sum of squares sum_nz (x~ - mu_c)^2 for the total-variance
// guard (cpu.rs:1130-1171): per-block f64 partials, fixed-order reduction.
__global__ void __launch_bounds__(256) k_centered_ss(const int64_t *__restrict__ indptr,
                                                     const uint32_t *__restrict__ idx,
                                                     const float *__restrict__ val,
                                                     const double *__restrict__ rowf, int flags,
                                                     const double *__restrict__ mu, int64_t n_obs,
                                                     double *__restrict__ part) {
    __shared__ double red[256];
    const int lane = threadIdx.x & 63;
    double acc = 0.0;
    const int64_t wstride = static_cast<int64_t>(gridDim.x) * (blockDim.x / kWave);
    for (int64_t r = static_cast<int64_t>(blockIdx.x) * (blockDim.x / kWave) + threadIdx.x / kWave;
         r < n_obs; r += wstride) {
        const int64_t a = indptr[r], b = indptr[r + 1];
        const double f = rowf ? rowf[r] : 1.0;
        for (int64_t p = a + lane; p < b; p += kWave) {
            const double d = static_cast<double>(xtilde(val[p], f, flags)) - mu[idx[p]];
            acc += d * d;
        }
    }
    red[threadIdx.x] = acc;
    __syncthreads();
    if (threadIdx.x == 0) {
        double s = 0.0;
        for (int i = 0; i < static_cast<int>(blockDim.x); ++i) s += red[i];
        part[blockIdx.x] = s;
    }
}

// ---------------------------------------------------------------------------
// Weighted column sums: part[blk][j] = sum_{r in blk} w[r] * A[r][j]
// (w == nullptr -> 1). mc = mu^T M (cpu.rs:1198-1206) and 1^T Q (cpu.rs:1238).
// ---------------------------------------------------------------------------

template <int KP, typename T>
__global__ void __launch_bounds__(256) k_colsum(const T *__restrict__ A, const double *__restrict__ w,
                                                int64_t nrows, int64_t rows_per_blk,
                                                double *__restrict__ part) {
    constexpr int RL = 256 / KP;
    __shared__ double red[256];
    const int j = threadIdx.x % KP, rl = threadIdx.x / KP;
    const int64_t r0 = static_cast<int64_t>(blockIdx.x) * rows_per_blk;
    const int64_t r1 = r0 + rows_per_blk < nrows ? r0 + rows_per_blk : nrows;
    double acc = 0.0;
    for (int64_t r = r0 + rl; r < r1; r += RL) {
        const double a = static_cast<double>(A[r * KP + j]);
        acc += w ? w[r] * a : a;
    }
    red[threadIdx.x] = acc;
    __syncthreads();
    if (rl == 0) {
        double s = 0.0;
        for (int i = 0; i < RL; ++i) s += red[i * KP + j];
        part[static_cast<int64_t>(blockIdx.x) * KP + j] = s;
    }
}

template <int KP>
__global__ void k_colsum_reduce(const double *__restrict__ part, int nblk, double *__restrict__ out64,
                                float *__restrict__ out32) {
    const int j = threadIdx.x;
    if (j >= KP) return;
    double s = 0.0;
    for (int b = 0; b < nblk; ++b) s += part[static_cast<int64_t>(b) * KP + j];
    if (out64) out64[j] = s;
    if (out32) out32[j] = static_cast<float>(s);
}

// ---------------------------------------------------------------------------
// Forward SpMM: Y[r][:] = sum_{nz in row r} x~ * M[col][:] - mc   (cpu.rs:1471-1523)
// One wavefront per row; a batch of 64 nonzeros is transformed lane-parallel
// and staged in LDS; the four 16-lane groups then take nonzeros 4t+g and
// gather M rows (L2-resident panel) as float4; the groups' partial rows are
// summed with cross-lane shuffles.
// ---------------------------------------------------------------------------

template <int KP, bool BF16>
__global__ void __launch_bounds__(256) k_spmm_fwd(const int64_t *__restrict__ indptr,
                                                  const uint32_t *__restrict__ idx,
                                                  const float *__restrict__ val,
                                                  const uint16_t *__restrict__ xbf,
                                                  const double *__restrict__ rowf, int flags,
                                                  const float *__restrict__ M, const float *__restrict__ mc,
                                                  float *__restrict__ Y, int64_t n_obs) {
    constexpr int E = KP / 16;
    __shared__ uint2 stage[4][kWave];
    const int wave = threadIdx.x / kWave, lane = threadIdx.x & 63, g = lane >> 4, l = lane & 15;
    const int64_t wstride = static_cast<int64_t>(gridDim.x) * 4;
    for (int64_t r = static_cast<int64_t>(blockIdx.x) * 4 + wave; r < n_obs; r += wstride) {
        const int64_t a = indptr[r], b = indptr[r + 1];
        const double f = (!BF16 && rowf) ? rowf[r] : 1.0;
        float acc[E];
#pragma unroll
        for (int e = 0; e < E; ++e) acc[e] = 0.f;
        for (int64_t base = a; base < b; base += kWave) {
            const int cnt = (b - base) < kWave ? static_cast<int>(b - base) : kWave;
            uint32_t c = 0u;
            float x = 0.f;
            if (lane < cnt) {
                c = idx[base + lane];
                x = BF16 ? bf16_to_f32(xbf[base + lane]) : xtilde(val[base + lane], f, flags);
            }
            stage[wave][lane] = make_uint2(c, __float_as_uint(x));
            wave_lds_sync();
            const int steps = (cnt + 3) >> 2;
            for (int t = 0; t < steps; ++t) {
                const uint2 ent = stage[wave][4 * t + g];
                const float xv = __uint_as_float(ent.y);
                const float4 *mr = reinterpret_cast<const float4 *>(M + static_cast<int64_t>(ent.x) * KP + l * E);
#pragma unroll
                for (int e = 0; e < E / 4; ++e) {
                    const float4 m4 = mr[e];
                    acc[4 * e + 0] += xv * m4.x;
                    acc[4 * e + 1] += xv * m4.y;
                    acc[4 * e + 2] += xv * m4.z;
                    acc[4 * e + 3] += xv * m4.w;
                }
            }
            wave_lds_sync();
        }
#pragma unroll
        for (int e = 0; e < E; ++e) {
            acc[e] += __shfl_xor(acc[e], 16);
            acc[e] += __shfl_xor(acc[e], 32);
        }
        if (g == 0) {
            float4 *yr = reinterpret_cast<float4 *>(Y + r * KP + l * E);
#pragma unroll
            for (int e = 0; e < E / 4; ++e) {
                float4 o;
                o.x = acc[4 * e + 0] - (mc ? mc[l * E + 4 * e + 0] : 0.f);
                o.y = acc[4 * e + 1] - (mc ? mc[l * E + 4 * e + 1] : 0.f);
                o.z = acc[4 * e + 2] - (mc ? mc[l * E + 4 * e + 2] : 0.f);
                o.w = acc[4 * e + 3] - (mc ? mc[l * E + 4 * e + 3] : 0.f);
                yr[e] = o;
            }
        }
    }
}

// ---------------------------------------------------------------------------
// Transposed SpMM partials: part[rb][v][:] = sum_{r in rb} x~_rv * Q[r][:]
// (cpu.rs:1282-1389). Block (rb, gc) keeps the f32 accumulator of genes
// [gc*CG, gc*CG+CG) in LDS (CG*KP*4 = 128 KB); rows are split at gene-chunk
// boundaries by the precomputed spl offsets, so each nonzero is read once.
// Blocks of one row block are mapped to one XCD, dispatched back to back, so
// they share Q rows and row offsets through that XCD's L2.
// ---------------------------------------------------------------------------

template <int KP, int CG, bool BF16>
__global__ void __launch_bounds__(512) k_spmm_t(const int64_t *__restrict__ indptr,
                                                const uint32_t *__restrict__ idx,
                                                const float *__restrict__ val,
                                                const uint16_t *__restrict__ xbf,
                                                const double *__restrict__ rowf, int flags,
                                                const float *__restrict__ Q, const uint32_t *__restrict__ spl,
                                                int nsplit, int cpc, int nch, int nrb, int64_t n_obs,
                                                int64_t n_vars, int64_t rows_per_rb,
                                                float *__restrict__ part) {
    constexpr int E = KP / 16;
    __shared__ float zs[CG * KP];
    __shared__ uint2 stage[8][kWave];
    // XCD-aware block mapping: blockIdx % 8 is the XCD; consecutive slots on one
    // XCD walk the gene chunks of one row block.
    const int xcd = blockIdx.x & 7, slot = blockIdx.x >> 3;
    const int rb = (slot / nch) * 8 + xcd, gc = slot % nch;
    if (rb >= nrb) return;
    const int64_t g0 = static_cast<int64_t>(gc) * CG;
    for (int i = threadIdx.x; i < CG * KP; i += blockDim.x) zs[i] = 0.f;
    __syncthreads();
    const int wave = threadIdx.x / kWave, lane = threadIdx.x & 63, g = lane >> 4, l = lane & 15;
    const int nw = blockDim.x / kWave;
    const int64_t r0 = static_cast<int64_t>(rb) * rows_per_rb;
    const int64_t r1 = r0 + rows_per_rb < n_obs ? r0 + rows_per_rb : n_obs;
    const int s_lo = gc * cpc, s_hi = (gc + 1) * cpc < nsplit ? (gc + 1) * cpc : nsplit;
    for (int64_t r = r0 + wave; r < r1; r += nw) {
        const int64_t a = indptr[r];
        const int64_t lo = a + spl[r * (nsplit + 1) + s_lo], hi = a + spl[r * (nsplit + 1) + s_hi];
        if (lo >= hi) continue;
        const double f = (!BF16 && rowf) ? rowf[r] : 1.0;
        float qv[E];
        const float4 *qr = reinterpret_cast<const float4 *>(Q + r * KP + l * E);
#pragma unroll
        for (int e = 0; e < E / 4; ++e) {
            const float4 q4 = qr[e];
            qv[4 * e] = q4.x; qv[4 * e + 1] = q4.y; qv[4 * e + 2] = q4.z; qv[4 * e + 3] = q4.w;
        }
        for (int64_t base = lo; base < hi; base += kWave) {
            const int cnt = (hi - base) < kWave ? static_cast<int>(hi - base) : kWave;
            uint32_t c = 0u;
            float x = 0.f;
            if (lane < cnt) {
                c = idx[base + lane] - static_cast<uint32_t>(g0);
                x = BF16 ? bf16_to_f32(xbf[base + lane]) : xtilde(val[base + lane], f, flags);
            }
            stage[wave][lane] = make_uint2(c, __float_as_uint(x));
            wave_lds_sync();
            const int steps = (cnt + 3) >> 2;
            for (int t = 0; t < steps; ++t) {
                const int s = 4 * t + g;
                if (s < cnt) {
                    const uint2 ent = stage[wave][s];
                    const float xv = __uint_as_float(ent.y);
                    float *zr = zs + ent.x * KP + l * E;
#pragma unroll
                    for (int e = 0; e < E; ++e) atomicAdd(zr + e, xv * qv[e]);
                }
            }
            wave_lds_sync();
        }
    }
    __syncthreads();
    const int64_t ng = (g0 + CG <= n_vars) ? CG : (n_vars - g0);
    float *dst = part + (static_cast<int64_t>(rb) * n_vars + g0) * KP;
    for (int64_t i = threadIdx.x; i < ng * KP; i += blockDim.x) dst[i] = zs[i];
}

// B[v][j] = sum_rb part[rb][v][j] - mu[v] * sumq[j]  (f64; cpu.rs:1253-1267)
template <int KP>
__global__ void __launch_bounds__(256) k_spmm_t_reduce(const float *__restrict__ part, int nrb,
                                                       int64_t n_vars, const double *__restrict__ mu,
                                                       const double *__restrict__ sumq,
                                                       double *__restrict__ B, float *__restrict__ B32) {
    const int64_t t = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (t >= n_vars * KP) return;
    const int64_t v = t / KP;
    const int j = static_cast<int>(t % KP);
    double s = 0.0;
    for (int rb = 0; rb < nrb; ++rb) s += static_cast<double>(part[static_cast<int64_t>(rb) * n_vars * KP + t]);
    if (mu) s -= mu[v] * sumq[j];
    if (B) B[t] = s;
    if (B32) B32[t] = static_cast<float>(s);
}

// ---------------------------------------------------------------------------
// Tall-skinny Gram G = Y^T Y on MFMA (v_mfma_f32_32x32x2f32, FP32 accumulate).
// Each wavefront accumulates the upper-triangular 32x32 tiles over its row
// range and writes them as an f64 partial; partials are summed in f64 in a
// fixed order (k_gram_reduce), so the result is deterministic.
//   operand layout: lane l supplies A[l%32][l/32] and B[l/32][l%32];
//   accumulator reg q of lane l holds D[8(q/4) + 4(l/32) + q%4][l%32].
// ---------------------------------------------------------------------------

template <int KP>
__global__ void __launch_bounds__(256) k_gram(const float *__restrict__ Y, int64_t nrows,
                                              int64_t rows_per_wave, double *__restrict__ part) {
    constexpr int NT = KP / 32;
    constexpr int NTILE = NT * (NT + 1) / 2;
    const int lane = threadIdx.x & 63;
    const int64_t w = static_cast<int64_t>(blockIdx.x) * (blockDim.x / kWave) + threadIdx.x / kWave;
    const int64_t r0 = w * rows_per_wave;
    const int64_t r1 = r0 + rows_per_wave < nrows ? r0 + rows_per_wave : nrows;
    v16f acc[NTILE];
#pragma unroll
    for (int t = 0; t < NTILE; ++t) acc[t] = v16f{};
    const int kr = lane >> 5, cl = lane & 31;
    for (int64_t r = r0; r < r1; r += 2) {
        const int64_t rr = r + kr;
        float v[NT];
#pragma unroll
        for (int t = 0; t < NT; ++t) v[t] = rr < r1 ? Y[rr * KP + t * 32 + cl] : 0.f;
        int tt = 0;
#pragma unroll
        for (int ti = 0; ti < NT; ++ti)
#pragma unroll
            for (int tj = ti; tj < NT; ++tj) {
                acc[tt] = __builtin_amdgcn_mfma_f32_32x32x2f32(v[ti], v[tj], acc[tt], 0, 0, 0);
                ++tt;
            }
    }
    double *dst = part + w * KP * KP;
    int tt = 0;
#pragma unroll
    for (int ti = 0; ti < NT; ++ti)
#pragma unroll
        for (int tj = ti; tj < NT; ++tj) {
#pragma unroll
            for (int q = 0; q < 16; ++q) {
                const int i = ti * 32 + 8 * (q / 4) + 4 * kr + (q % 4);
                const int j = tj * 32 + cl;
                dst[i * KP + j] = static_cast<double>(acc[tt][q]);
            }
            ++tt;
        }
}

// G[i][j] (full symmetric) from the upper-tile partials.
template <int KP>
__global__ void __launch_bounds__(256) k_gram_reduce(const double *__restrict__ part, int nparts,
                                                     double *__restrict__ G) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= KP * KP) return;
    int i = t / KP, j = t % KP;
    if (i / 32 > j / 32) { const int s = i; i = j; j = s; }
    double s = 0.0;
    for (int p = 0; p < nparts; ++p) s += part[static_cast<int64_t>(p) * KP * KP + i * KP + j];
    G[t] = s;
}

// Gram of a small f64 matrix (B: n_vars x KP) on the VALU in f64.
template <int KP>
__global__ void __launch_bounds__(256) k_gram64(const double *__restrict__ A, int64_t nrows,
                                                int64_t rows_per_blk, double *__restrict__ part) {
    constexpr int PER = KP * KP / 256;
    __shared__ double tile[16 * KP];
    double acc[PER];
#pragma unroll
    for (int e = 0; e < PER; ++e) acc[e] = 0.0;
    const int64_t r0 = static_cast<int64_t>(blockIdx.x) * rows_per_blk;
    const int64_t r1 = r0 + rows_per_blk < nrows ? r0 + rows_per_blk : nrows;
    for (int64_t rb = r0; rb < r1; rb += 16) {
        __syncthreads();
        for (int i = threadIdx.x; i < 16 * KP; i += blockDim.x) {
            const int64_t r = rb + i / KP;
            tile[i] = r < r1 ? A[r * KP + (i % KP)] : 0.0;
        }
        __syncthreads();
        for (int rr = 0; rr < 16; ++rr) {
#pragma unroll
            for (int e = 0; e < PER; ++e) {
                const int o = threadIdx.x + e * 256;
                acc[e] += tile[rr * KP + o / KP] * tile[rr * KP + o % KP];
            }
        }
    }
#pragma unroll
    for (int e = 0; e < PER; ++e) part[static_cast<int64_t>(blockIdx.x) * KP * KP + threadIdx.x + e * 256] = acc[e];
}

template <int KP>
__global__ void __launch_bounds__(256) k_sum_parts(const double *__restrict__ part, int nparts,
                                                   double *__restrict__ G) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= KP * KP) return;
    double s = 0.0;
    for (int p = 0; p < nparts; ++p) s += part[static_cast<int64_t>(p) * KP * KP + t];
    G[t] = s;
}

// ---------------------------------------------------------------------------
// Cholesky (f64, one workgroup) of the leading k x k block of G (+ shift*trace
// on the diagonal when shift > 0): G = R^T R, R upper. Writes R (f64), its
// inverse W = R^-1 (f64 and f32, zero outside k x k) and a status word
// (1 = not SPD, reported like cusolver.rs:381-385).
// ---------------------------------------------------------------------------

template <int KP>
__global__ void __launch_bounds__(256) k_chol(const double *__restrict__ G, int k, double shift,
                                              double *__restrict__ R64, double *__restrict__ W64,
                                              float *__restrict__ W32, int *__restrict__ status) {
    __shared__ double a[KP * (KP + 1)];
    __shared__ int bad;
    constexpr int LD = KP + 1;
    const int tid = threadIdx.x;
    if (tid == 0) bad = 0;
    for (int t = tid; t < KP * KP; t += blockDim.x) a[(t / KP) * LD + t % KP] = G[t];
    __syncthreads();
    if (shift > 0.0) {
        __shared__ double tr;
        if (tid == 0) {
            double s = 0.0;
            for (int i = 0; i < k; ++i) s += a[i * LD + i];
            tr = s;
        }
        __syncthreads();
        for (int i = tid; i < k; i += blockDim.x) a[i * LD + i] += shift * tr;
        __syncthreads();
    }
    for (int j = 0; j < k; ++j) {
        if (tid == 0) {
            const double d = a[j * LD + j];
            if (!(d > 0.0) || !isfinite(d)) bad = 1;
            a[j * LD + j] = (d > 0.0 && isfinite(d)) ? sqrt(d) : 1.0;
        }
        __syncthreads();
        const double rjj = a[j * LD + j];
        for (int i = j + 1 + tid; i < k; i += blockDim.x) a[j * LD + i] /= rjj;
        __syncthreads();
        const int m = k - j - 1;
        for (int t = tid; t < m * m; t += blockDim.x) {
            const int i = j + 1 + t / m, c = j + 1 + t % m;
            if (c >= i) a[i * LD + c] -= a[j * LD + i] * a[j * LD + c];
        }
        __syncthreads();
    }
    // R (upper), zero elsewhere
    for (int t = tid; t < KP * KP; t += blockDim.x) {
        const int i = t / KP, c = t % KP;
        R64[t] = (i < k && c < k && c >= i) ? a[i * LD + c] : 0.0;
    }
    __syncthreads();
    // W = R^-1 by back substitution, one column per thread.
    for (int c = tid; c < KP; c += blockDim.x) {
        for (int i = KP - 1; i >= 0; --i) {
            double wv = 0.0;
            if (c < k && i <= c) {
                double s = (i == c) ? 1.0 : 0.0;
                for (int q = i + 1; q <= c; ++q) s -= a[i * LD + q] * W64[q * KP + c];
                wv = s / a[i * LD + i];
            }
            W64[i * KP + c] = wv;
        }
    }
    __syncthreads();
    for (int t = tid; t < KP * KP; t += blockDim.x) W32[t] = static_cast<float>(W64[t]);
    if (tid == 0 && bad) atomicOr(status, 1);
}

// Y <- Y W for a tall panel (KP x KP W in LDS) on MFMA; writes out_cols
// columns of the product with row stride out_ld. The 32-row tile of Y is staged
// in LDS with a padded stride so the A operand reads are conflict free.
template <int KP>
__global__ void __launch_bounds__(256) k_apply(const float *__restrict__ Y, const float *__restrict__ W,
                                               float *__restrict__ out, int64_t nrows, int out_cols,
                                               int64_t out_ld) {
    constexpr int NT = KP / 32;
    constexpr int LDY = KP + 1;
    __shared__ float ws[KP * KP];
    __shared__ float ys[4][32 * LDY];
    for (int t = threadIdx.x; t < KP * KP; t += blockDim.x) ws[t] = W[t];
    __syncthreads();
    const int wave = threadIdx.x / kWave, lane = threadIdx.x & 63;
    const int kr = lane >> 5, cl = lane & 31;
    const int64_t ntiles = (nrows + 31) / 32;
    for (int64_t tile = static_cast<int64_t>(blockIdx.x) * 4 + wave; tile < ntiles;
         tile += static_cast<int64_t>(gridDim.x) * 4) {
        const int64_t r0 = tile * 32;
        float *yt = ys[wave];
        for (int i = lane; i < 32 * KP; i += kWave) {
            const int rr = i / KP, cc = i % KP;
            yt[rr * LDY + cc] = (r0 + rr < nrows) ? Y[(r0 + rr) * KP + cc] : 0.f;
        }
        wave_lds_sync();
#pragma unroll
        for (int nt = 0; nt < NT; ++nt) {
            v16f acc = v16f{};
            for (int s = 0; s < KP / 2; ++s) {
                const float av = yt[cl * LDY + 2 * s + kr];
                const float bv = ws[(2 * s + kr) * KP + nt * 32 + cl];
                acc = __builtin_amdgcn_mfma_f32_32x32x2f32(av, bv, acc, 0, 0, 0);
            }
            const int col = nt * 32 + cl;
            if (col < out_cols) {
#pragma unroll
                for (int q = 0; q < 16; ++q) {
                    const int64_t row = r0 + 8 * (q / 4) + 4 * kr + (q % 4);
                    if (row < nrows) out[row * out_ld + col] = acc[q];
                }
            }
        }
        wave_lds_sync();
    }
}

// out = A W for a small f64 A (rows x KP), W f64 KP x KP; thread per element.
template <int KP>
__global__ void __launch_bounds__(256) k_apply64(const double *__restrict__ A, const double *__restrict__ W,
                                                 double *__restrict__ out, float *__restrict__ out32,
                                                 int64_t nrows) {
    const int64_t t = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (t >= nrows * KP) return;
    const int64_t r = t / KP;
    const int j = static_cast<int>(t % KP);
    double s = 0.0;
    for (int i = 0; i <= j; ++i) s += A[r * KP + i] * W[i * KP + j];  // W upper triangular or dense below
    for (int i = j + 1; i < KP; ++i) s += A[r * KP + i] * W[i * KP + j];
    if (out) out[t] = s;
    if (out32) out32[t] = static_cast<float>(s);
}

// C = A B for KP x KP f64 matrices (one workgroup).
template <int KP>
__global__ void __launch_bounds__(256) k_small_gemm64(const double *__restrict__ A, const double *__restrict__ B,
                                                      double *__restrict__ C) {
    for (int t = threadIdx.x; t < KP * KP; t += blockDim.x) {
        const int i = t / KP, j = t % KP;
        double s = 0.0;
        for (int q = 0; q < KP; ++q) s += A[i * KP + q] * B[q * KP + j];
        C[t] = s;
    }
}

// ---------------------------------------------------------------------------
// One-sided (Hestenes) Jacobi SVD of the k x k upper-triangular R in f64,
// one workgroup, round-robin (circle method) pair ordering. Outputs sigma
// (descending), U (columns A_j / sigma_j) and V, all KP x KP row-major with
// columns permuted to the descending order.
// ---------------------------------------------------------------------------

template <int KP>
__global__ void __launch_bounds__(256) k_jacobi_svd(const double *__restrict__ R, int k,
                                                    double *__restrict__ Acol, double *__restrict__ Vcol,
                                                    double *__restrict__ U, double *__restrict__ V,
                                                    double *__restrict__ sigma, int *__restrict__ status) {
    // Acol / Vcol: column-major KP x KP scratch in global memory (L2 resident).
    __shared__ int changed;
    __shared__ double nrm[KP];
    __shared__ int order[KP];
    const int tid = threadIdx.x;
    const int n = (k + 1) & ~1;  // even count; the extra column (if any) is zero
    for (int t = tid; t < KP * KP; t += blockDim.x) {
        const int i = t % KP, c = t / KP;  // column-major index t = c*KP + i
        Acol[t] = (i < k && c < k) ? R[i * KP + c] : 0.0;
        Vcol[t] = (i == c && c < n) ? 1.0 : 0.0;
    }
    __syncthreads();
    constexpr int TPP = 8;  // threads per pair
    const int npairs = n / 2;
    for (int sweep = 0; sweep < 60 && n >= 2; ++sweep) {
        if (tid == 0) changed = 0;
        __syncthreads();
        for (int round = 0; round < n - 1; ++round) {
            for (int pb = tid / TPP; pb < npairs; pb += blockDim.x / TPP) {
                int p, q;
                if (pb == 0) {
                    p = 0;
                    q = 1 + (round % (n - 1));
                } else {
                    p = 1 + (round + pb) % (n - 1);
                    q = 1 + (round - pb + (n - 1)) % (n - 1);
                }
                if (p > q) { const int s = p; p = q; q = s; }
                double *ap = Acol + static_cast<int64_t>(p) * KP, *aq = Acol + static_cast<int64_t>(q) * KP;
                const int sub = tid % TPP;
                double al = 0.0, be = 0.0, ga = 0.0;
                for (int i = sub; i < n; i += TPP) {
                    const double x = ap[i], y = aq[i];
                    al += x * x; be += y * y; ga += x * y;
                }
                for (int o = TPP / 2; o > 0; o >>= 1) {
                    al += __shfl_xor(al, o, TPP);
                    be += __shfl_xor(be, o, TPP);
                    ga += __shfl_xor(ga, o, TPP);
                }
                if (ga != 0.0 && fabs(ga) > 1e-15 * sqrt(al * be)) {
                    const double zeta = (be - al) / (2.0 * ga);
                    const double tt = (zeta >= 0 ? 1.0 : -1.0) / (fabs(zeta) + sqrt(1.0 + zeta * zeta));
                    const double c = 1.0 / sqrt(1.0 + tt * tt), s = c * tt;
                    for (int i = sub; i < n; i += TPP) {
                        const double x = ap[i], y = aq[i];
                        ap[i] = c * x - s * y;
                        aq[i] = s * x + c * y;
                    }
                    double *vp = Vcol + static_cast<int64_t>(p) * KP, *vq = Vcol + static_cast<int64_t>(q) * KP;
                    for (int i = sub; i < n; i += TPP) {
                        const double x = vp[i], y = vq[i];
                        vp[i] = c * x - s * y;
                        vq[i] = s * x + c * y;
                    }
                    if (sub == 0) changed = 1;
                }
            }
            __threadfence_block();
            __syncthreads();
        }
        if (!changed) break;
        if (sweep == 59 && tid == 0) atomicOr(status, 2);
        __syncthreads();
    }
    for (int c = tid; c < KP; c += blockDim.x) {
        double s = 0.0;
        if (c < k)
            for (int i = 0; i < n; ++i) s += Acol[static_cast<int64_t>(c) * KP + i] * Acol[static_cast<int64_t>(c) * KP + i];
        nrm[c] = c < k ? sqrt(s) : -1.0;
    }
    __syncthreads();
    // rank by descending norm (ties by index) -> stable order
    for (int c = tid; c < KP; c += blockDim.x) {
        int rank = 0;
        for (int d = 0; d < KP; ++d)
            if (nrm[d] > nrm[c] || (nrm[d] == nrm[c] && d < c)) ++rank;
        order[rank] = c;
    }
    __syncthreads();
    for (int t = tid; t < KP * KP; t += blockDim.x) {
        const int i = t / KP, jj = t % KP;  // output row-major (i, jj)
        const int c = order[jj];
        const double sg = nrm[c] > 0.0 ? nrm[c] : 0.0;
        U[t] = (jj < k && sg > 0.0) ? Acol[static_cast<int64_t>(c) * KP + i] / sg : 0.0;
        V[t] = jj < k ? Vcol[static_cast<int64_t>(c) * KP + i] : 0.0;
    }
    for (int jj = tid; jj < KP; jj += blockDim.x) sigma[jj] = jj < k ? (nrm[order[jj]] > 0 ? nrm[order[jj]] : 0.0) : 0.0;
}

// Embedding weights W_emb = (V Sigma)[:, :nc] as f32 KP x KP (zero padded).
template <int KP>
__global__ void k_emb_weights(const double *__restrict__ V, const double *__restrict__ sigma, int nc,
                              float *__restrict__ W) {
    for (int t = threadIdx.x; t < KP * KP; t += blockDim.x) {
        const int j = t % KP;
        W[t] = j < nc ? static_cast<float>(V[t] * sigma[j]) : 0.f;
    }
}

// components[pc][v] = Uhat[v][pc] for pc < nc (f32).
template <int KP>
__global__ void __launch_bounds__(256) k_components(const double *__restrict__ Uhat, int nc, int64_t n_vars,
                                                    float *__restrict__ comps) {
    const int64_t t = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (t >= n_vars * nc) return;
    const int64_t pc = t / n_vars, v = t % n_vars;
    comps[t] = static_cast<float>(Uhat[v * KP + pc]);
}

// Panel M (f32 n_vars x KP) from a row-major f64 (n_vars x k) host matrix copy.
template <int KP>
__global__ void __launch_bounds__(256) k_pack_panel(const double *__restrict__ src, int64_t rows, int k,
                                                    float *__restrict__ dst) {
    const int64_t t = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (t >= rows * KP) return;
    const int64_t r = t / KP;
    const int j = static_cast<int>(t % KP);
    dst[t] = j < k ? static_cast<float>(src[r * k + j]) : 0.f;
}

template <int KP>
__global__ void __launch_bounds__(256) k_pack_panel32(const float *__restrict__ src, int64_t rows, int k,
                                                      float *__restrict__ dst) {
    const int64_t t = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (t >= rows * KP) return;
    const int64_t r = t / KP;
    const int j = static_cast<int>(t % KP);
    dst[t] = j < k ? src[r * k + j] : 0.f;
}

template <int KP>
__global__ void __launch_bounds__(256) k_unpack_panel32(const float *__restrict__ src, int64_t rows, int k,
                                                        float *__restrict__ dst) {
    const int64_t t = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (t >= rows * k) return;
    const int64_t r = t / k;
    const int j = static_cast<int>(t % k);
    dst[t] = src[r * KP + j];
}

__global__ void k_f32_to_f64(const float *__restrict__ s, double *__restrict__ d, int64_t n) {
    const int64_t t = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (t < n) d[t] = static_cast<double>(s[t]);
}

__global__ void k_f64_to_f32(const double *__restrict__ s, float *__restrict__ d, int64_t n) {
    const int64_t t = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (t < n) d[t] = static_cast<float>(s[t]);
}

}  // namespace scxpca
