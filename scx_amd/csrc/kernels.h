// HIP kernels (gfx950) for the randomized sparse PCA.
//
// Data layout in HBM:
//   CSR          indptr i64[n_obs+1], indices u32[nnz], values f32[nnz] (raw counts)
//   rowf         f64[n_obs]   normalize_total factor target/rowsum (1 where rowsum <= 0)
//   xt           f32[nnz]     X~ = log1p(f32(x * rowf)) in CSR order, computed once per
//                             prepared matrix (read by the FP32 forward product)
//   spl          u32[n_obs][nsplit+1]  offset of the first nonzero of each CB-gene
//                chunk inside each row (sorted rows)
//   tile-CSC     per 256-row tile, gene-major runs: tptr u32[n_tiles][n_vars+1],
//                trow u8[nnz] (row in tile), tval f32[nnz] (X~); tpk u32[nnz] packed
//                bf16 entries for the BF16 sketch form
//   xpk          u32[nnz] packed bf16 entries in CSR order (BF16 sketch form)
//   panels       f32 row-major, KP columns (k' padded to 32/64/128, pad columns 0):
//                M (n_vars x KP), Y and Q (n_obs x KP)
//   B            f64 n_vars x KP
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <type_traits>

namespace scxpca {

constexpr int kWave = 64;
constexpr int kXfNorm = 1;
constexpr int kXfLog1p = 2;
// PFlog (v4): every row's factor is the matrix-wide 4 alpha (no row totals);
// kXfLogF64 takes the log in f64 and rounds once (the in-memory delta builder);
// kXfCheck rejects non-finite values (and, with kXfLogF64, negative counts).
constexpr int kXfScale = 4;
constexpr int kXfLogF64 = 8;
constexpr int kXfCheck = 16;

typedef float v16f __attribute__((ext_vector_type(16)));

__device__ __forceinline__ void wave_lds_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// log1p(x) in f32 without the double-float arithmetic of the library routine:
// u = 1 + x rounded, the rounding error folded back as c/u, u reduced to
// [sqrt(2)/2, sqrt(2)) by exponent arithmetic, and log(1+f) = 2 atanh(f/(2+f))
// by the 4-term polynomial of FreeBSD/musl logf (Sun fdlibm, public domain;
// |error of the polynomial| < 2^-34.24). Non-finite inputs and x <= -1 take
// the library routine.
__host__ __device__ __forceinline__ float rcp_approx(float v) {
#ifdef __HIP_DEVICE_COMPILE__
    return __builtin_amdgcn_rcpf(v);
#else
    return 1.0f / v;
#endif
}

__host__ __device__ __forceinline__ float log1p_fast(float x) {
    if (!(x > -1.0f && x < 3.0e37f)) return log1pf(x);
    const float ln2_hi = 6.9313812256e-01f, ln2_lo = 9.0580006145e-06f;
    const float Lg1 = 0.66666662693f, Lg2 = 0.40000972152f, Lg3 = 0.28498786688f, Lg4 = 0.24279078841f;
    const float u = 1.0f + x;
    uint32_t iu = __builtin_bit_cast(uint32_t, u) + (0x3f800000u - 0x3f3504f3u);
    const int k = static_cast<int>(iu >> 23) - 0x7f;
    float c = (k >= 2) ? 1.0f - (u - x) : x - (u - 1.0f);
    c = (k < 25) ? c * rcp_approx(u) : 0.0f;
    iu = (iu & 0x007fffffu) + 0x3f3504f3u;
    const float f = __builtin_bit_cast(float, iu) - 1.0f;
    const float s = f * rcp_approx(2.0f + f);
    const float z = s * s, w = z * z;
    const float R = z * (Lg1 + w * Lg3) + w * (Lg2 + w * Lg4);
    const float hfsq = 0.5f * f * f;
    const float dk = static_cast<float>(k);
    return s * (hfsq + R) + (dk * ln2_lo + c) - hfsq + f + dk * ln2_hi;
}

// transforms.rs:21-39: ((x as f64 * factor) as f32).ln_1p(); the factor is 1
// for rows whose total is not positive, which reproduces the :40-72 branch.
// PFlog: the Scale{4a} -> Log1p chain (transforms.rs:109-128) is the same
// expression with f = 4a; the in-memory builder (pyscx/src/accel/pflog.rs:
// 375-397) is (4a * x as f64).ln_1p() as f32.
template <bool LOGF64>  // LOGF64: the in-memory PFlog builder (kXfLogF64)
__device__ __forceinline__ float xtilde(float v, double f, int flags) {
    if constexpr (LOGF64) return static_cast<float>(log1p(f * static_cast<double>(v)));
    const float s = (flags & (kXfNorm | kXfScale)) ? static_cast<float>(static_cast<double>(v) * f) : v;
    return (flags & kXfLog1p) ? log1p_fast(s) : s;
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

// The BF16 sketch form's packed tile-CSC entry: bf16(x~) (bits 0-15), row in
// the 256-row tile (16-23), gene mod 64 (24-29).
__device__ __forceinline__ uint32_t pack_bf16_entry(float x, uint32_t row, uint32_t gene) {
    return static_cast<uint32_t>(f32_to_bf16_rne(x)) | ((row & 255u) << 16) | ((gene & 63u) << 24);
}

// The BF16 sketch form's packed CSR-order entry: bf16(x~) (bits 16-31), the
// gene mod kChunkMod (bits 0-15): x~ is the entry with its low half cleared.
__device__ __forceinline__ uint32_t pack_bf16_row_entry(float x, uint32_t lgv) {
    return (static_cast<uint32_t>(f32_to_bf16_rne(x)) << 16) | (lgv & 0xffffu);
}

// x rounded to bf16 (round to nearest even) and back to f32
__device__ __forceinline__ float bf16_round(float x) { return bf16_to_f32(f32_to_bf16_rne(x)); }

template <typename T>
__device__ __forceinline__ T wave_sum(T v) {
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
    return v;
}

// ---------------------------------------------------------------------------
// One-wave f64 helpers: lane broadcast and a 64-lane sum by DPP (quad
// permutations and row mirrors inside each 16-lane row, then the four row
// sums read by readlane), the result identical in every lane.
// ---------------------------------------------------------------------------
__device__ __forceinline__ double readlane_f64(double v, int l) {
    const uint64_t u = __builtin_bit_cast(uint64_t, v);
    const uint32_t lo = static_cast<uint32_t>(__builtin_amdgcn_readlane(static_cast<int>(static_cast<uint32_t>(u)), l));
    const uint32_t hi = static_cast<uint32_t>(__builtin_amdgcn_readlane(static_cast<int>(static_cast<uint32_t>(u >> 32)), l));
    return __builtin_bit_cast(double, (static_cast<uint64_t>(hi) << 32) | lo);
}

template <int CTRL>
__device__ __forceinline__ double dpp_f64(double v) {
    const uint64_t u = __builtin_bit_cast(uint64_t, v);
    const uint32_t lo = static_cast<uint32_t>(
        __builtin_amdgcn_update_dpp(0, static_cast<int>(static_cast<uint32_t>(u)), CTRL, 0xf, 0xf, false));
    const uint32_t hi = static_cast<uint32_t>(
        __builtin_amdgcn_update_dpp(0, static_cast<int>(static_cast<uint32_t>(u >> 32)), CTRL, 0xf, 0xf, false));
    return __builtin_bit_cast(double, (static_cast<uint64_t>(hi) << 32) | lo);
}

// 1/sqrt(d) for finite d > 0: the hardware estimate (relative error about
// 2^-23) refined by two Newton-Raphson steps y += y (1/2 - d y^2 / 2), each
// doubling the correct bits. Branch-free, unlike the library routine.
__device__ __forceinline__ double rsq_nr(double d) {
    double y = __builtin_amdgcn_rsq(d);
    const double h = 0.5 * d;
    y = fma(y, fma(-h * y, y, 0.5), y);
    y = fma(y, fma(-h * y, y, 0.5), y);
    return y;
}

// Sum over each 16-lane DPP row, the same bits in all 16 lanes (every step
// adds the same two operands in every lane, in either order).
__device__ __forceinline__ double row16_sum_dpp(double v) {
    v += dpp_f64<0xB1>(v);   // quad_perm [1,0,3,2]
    v += dpp_f64<0x4E>(v);   // quad_perm [2,3,0,1]
    v += dpp_f64<0x141>(v);  // row_half_mirror
    v += dpp_f64<0x140>(v);  // row_mirror
    return v;
}

__device__ __forceinline__ double wave_sum_dpp(double v) {
    v = row16_sum_dpp(v);
    return (readlane_f64(v, 0) + readlane_f64(v, 16)) + (readlane_f64(v, 32) + readlane_f64(v, 48));
}


// ---------------------------------------------------------------------------
// Per-row preparation
// ---------------------------------------------------------------------------


// indptr check before anything dereferences entries through it: flag = 1
// when indptr[0] != 0 or any indptr[r + 1] < indptr[r].
__global__ void __launch_bounds__(256) k_check_indptr(const int64_t *__restrict__ indptr, int64_t n_obs,
                                                      int *__restrict__ flag) {
    const int64_t r = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (r >= n_obs) return;
    const bool bad = indptr[r + 1] < indptr[r] || (r == 0 && indptr[0] != 0);
    if (bad) atomicOr(flag, 1);
}



// The forward SpMM's LDS panel chunks hold 512 (KP = 64) or 256 (KP = 128)
// genes and start at multiples of their size, so gene mod 512 (stored as u16)
// gives the row inside either chunk (mod 256 for the smaller one).
constexpr uint32_t kChunkMod = 512;

// Per-row preparation in one pass over the CSR (one wave per row): the row
// total (lane partials, wave reduction; exact for integer counts) and factor
// t/s (1 where s <= 0; csr.rs:543-552, transforms.rs:21-72), X~ in CSR order,
// the chunk-local gene index lg and, when nsplit > 0,
// the split offsets spl[r][g] = number of entries with gene < g * 128 (rows
// are sorted: written by the entries that open a chunk, and the row length for
// chunks after the last entry; spl[r][nsplit] = row length).
// The same pass validates the matrix into *flag (bits: 1 a row's indices
// decrease, 2 an index >= n_vars, 4 indptr decreases, 8 a row repeats an
// index; with kXfCheck 16 a non-finite value or transformed value, 32 a
// negative count under kXfLogF64, the first such entry's position in
// ~*errpos); the tile-CSC builds skip all work when the flag is set, and the
// host rejects the matrix before anything else reads these outputs.
// PFlog rows (rowb != null) also get the baseline
// b_i = -(sum_j x~_ij) * (1 / n_vars) (pflog.rs:186-222), f64 lane partials
// and a wave reduction.
// PFLOG: the value checks (kXfCheck) and the baseline rowb; LOGF64: kXfLogF64
// (separate instantiations: the f64 log costs the plain pass registers)
template <bool PFLOG, bool LOGF64>
__global__ void __launch_bounds__(256) k_prepare_rows(const int64_t *__restrict__ indptr,
                                                      const uint32_t *__restrict__ idx,
                                                      const float *__restrict__ val, int64_t n_obs,
                                                      double target, int flags, int nsplit,
                                                      double *__restrict__ rowf, float *__restrict__ xt,
                                                      uint16_t *__restrict__ lg, uint32_t *__restrict__ xpk,
                                                      uint32_t *__restrict__ spl,
                                                      int64_t n_vars, int *__restrict__ flag,
                                                      double *__restrict__ rowb,
                                                      unsigned long long *__restrict__ errpos) {
    constexpr int U = 4;  // rows of up to U * 64 entries are loaded in one go
    const double inv_d = PFLOG ? 1.0 / static_cast<double>(n_vars > 0 ? n_vars : 1) : 0.0;
    const int lane = threadIdx.x & 63;
    const int64_t wstride = static_cast<int64_t>(gridDim.x) * (blockDim.x / kWave);
    int bad = 0;
    // One 64-entry batch of a row: the index checks against the previous entry
    // (lane 0: the previous batch's last, carried in prev_last; -1 before the
    // first), and the split offsets: an entry whose 128-gene chunk differs from
    // its predecessor's opens chunks (prev, ch]; their offsets are its position.
    auto scan = [&](uint32_t *sr, int64_t a, int64_t b, int64_t p0, bool ok, uint32_t c, int64_t &prev_last) {
        // lane - 1's index by a whole-wave DPP shift (wave_shr:1); lane 0
        // takes the previous batch's last
        const uint32_t up =
            static_cast<uint32_t>(__builtin_amdgcn_update_dpp(0, static_cast<int>(c), 0x138, 0xf, 0xf, false));
        const int64_t prev = lane == 0 ? prev_last : static_cast<int64_t>(up);
        if (ok) {
            bad |= (static_cast<int64_t>(c) >= n_vars) ? 2 : 0;
            bad |= (prev >= 0 && static_cast<int64_t>(c) < prev) ? 1 : 0;
            bad |= (static_cast<int64_t>(c) == prev) ? 8 : 0;
        }
        if (nsplit > 0 && ok) {
            const int ch = static_cast<int>(c >> 7), chp = prev >= 0 ? static_cast<int>(prev >> 7) : -1;
            if (ch > chp)
                for (int g = chp + 1; g <= ch && g < nsplit; ++g) sr[g] = static_cast<uint32_t>(p0 + lane - a);
        }
        const int cnt = (b - p0) < kWave ? static_cast<int>(b - p0) : kWave;  // wave-uniform
        prev_last = static_cast<int64_t>(static_cast<uint32_t>(__builtin_amdgcn_readlane(static_cast<int>(c), cnt - 1)));
    };
    auto splits_tail = [&](uint32_t *sr, int64_t a, int64_t b, int64_t prev_last) {
        // chunks after the row's last entry start at the row's end
        const int chl = prev_last >= 0 ? static_cast<int>(prev_last >> 7) : -1;
        for (int g = (chl + 1 > 0 ? chl + 1 : 0) + lane; g < nsplit; g += kWave) sr[g] = static_cast<uint32_t>(b - a);
        if (lane == 0) sr[nsplit] = static_cast<uint32_t>(b - a);
    };
    // kXfCheck: non-finite input or transformed value; a negative count for
    // the in-memory PFlog builder; the smallest offending position wins
    auto check_value = [&](float v, float d, int64_t p) {
        const bool nf = !__builtin_isfinite(v) || !__builtin_isfinite(d);
        const bool neg = (flags & kXfLogF64) && v < 0.f;
        if (nf || neg) {
            bad |= (nf ? 16 : 0) | (neg ? 32 : 0);
            atomicMax(errpos, ~static_cast<unsigned long long>(p));
        }
    };
    auto put_baseline = [&](int64_t r, double dsum) {
        dsum = wave_sum_dpp(dsum);
        if (lane == 0) rowb[r] = -dsum * inv_d;
    };
    auto factor = [&](int64_t r, double sum) {
        if constexpr (PFLOG)
            if (flags & kXfScale) return target;
        double f = 1.0;
        if (flags & kXfNorm) {
            sum = wave_sum_dpp(sum);
            f = sum > 0.0 ? target / sum : 1.0;
            if (lane == 0) rowf[r] = f;
        }
        return f;
    };
    // short rows (<= U * 64 entries): every load issued at once, clamped
    // into the row (no loads for an empty row)
    auto load_short = [&](int64_t a, int64_t b, uint32_t (&c)[U], float (&v)[U]) {
        if (b > a) {
#pragma unroll
            for (int u = 0; u < U; ++u) {
                int64_t p = a + u * kWave + lane;
                p = p < b - 1 ? p : b - 1;
                c[u] = idx[p];
                v[u] = val[p];
            }
        }
    };
    auto finish_short = [&](int64_t r, int64_t a, int64_t b, const uint32_t (&c)[U], const float (&v)[U]) {
        double sum = 0.0;
        if (flags & kXfNorm) {
#pragma unroll
            for (int u = 0; u < U; ++u)
                if (a + u * kWave + lane < b) sum += static_cast<double>(v[u]);
        }
        const double f = factor(r, sum);
        uint32_t *sr = nsplit > 0 ? spl + r * (nsplit + 1) : nullptr;
        int64_t prev_last = -1;
        double dsum = 0.0;
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const int64_t p0 = a + u * kWave;
            if (p0 >= b) break;
            const int64_t p = p0 + lane;
            const bool ok = p < b;
            if (ok) {
                const float d = xtilde<LOGF64>(v[u], f, flags);
                xt[p] = d;
                if (lg) lg[p] = static_cast<uint16_t>(c[u] & (kChunkMod - 1));
                if (xpk) xpk[p] = pack_bf16_row_entry(d, c[u] & (kChunkMod - 1));
                if constexpr (PFLOG) {
                    dsum += static_cast<double>(d);
                    if (flags & kXfCheck) check_value(v[u], d, p);
                }
            }
            scan(sr, a, b, p0, ok, c[u], prev_last);
        }
        if (nsplit > 0) splits_tail(sr, a, b, prev_last);
        if constexpr (PFLOG)
            if (rowb) put_baseline(r, dsum);
    };
    auto long_row = [&](int64_t r, int64_t a, int64_t b) {
        double sum = 0.0;
        if (flags & kXfNorm)
            for (int64_t p = a + lane; p < b; p += kWave) sum += static_cast<double>(val[p]);
        const double f = factor(r, sum);
        uint32_t *sr = nsplit > 0 ? spl + r * (nsplit + 1) : nullptr;
        int64_t prev_last = -1;
        double dsum = 0.0;
        for (int64_t p0 = a; p0 < b; p0 += kWave) {
            const int64_t p = p0 + lane;
            const bool ok = p < b;
            const uint32_t c = ok ? idx[p] : 0u;
            if (ok) {
                const float v = val[p];
                const float d = xtilde<LOGF64>(v, f, flags);
                xt[p] = d;
                if (lg) lg[p] = static_cast<uint16_t>(c & (kChunkMod - 1));
                if (xpk) xpk[p] = pack_bf16_row_entry(d, c & (kChunkMod - 1));
                if constexpr (PFLOG) {
                    dsum += static_cast<double>(d);
                    if (flags & kXfCheck) check_value(v, d, p);
                }
            }
            scan(sr, a, b, p0, ok, c, prev_last);
        }
        if (nsplit > 0) splits_tail(sr, a, b, prev_last);
        if constexpr (PFLOG)
            if (rowb) put_baseline(r, dsum);
    };
    // two rows per wave at a time (r and r + wstride): both rows' loads are in
    // flight together, so each wave keeps twice the bytes outstanding
    // the wave's index made wave-uniform: the row bounds and the per-row
    // bookkeeping then live in scalar registers (scalar loads of indptr)
    const int wib = __builtin_amdgcn_readfirstlane(static_cast<int>(threadIdx.x / kWave));
    for (int64_t r = static_cast<int64_t>(blockIdx.x) * (blockDim.x / kWave) + wib; r < n_obs; r += 2 * wstride) {
        const int64_t r2 = r + wstride;
        const bool has2 = r2 < n_obs;
        const int64_t a = indptr[r], b = indptr[r + 1];
        const int64_t a2 = has2 ? indptr[r2] : 0, b2 = has2 ? indptr[r2 + 1] : 0;
        bad |= (b < a || b2 < a2) ? 4 : 0;
        const bool s1 = b - a <= U * kWave, s2 = has2 && (b2 - a2 <= U * kWave);
        uint32_t c1[U] = {}, c2[U] = {};
        float v1[U] = {}, v2[U] = {};
        if (s1) load_short(a, b, c1, v1);
        if (s2) load_short(a2, b2, c2, v2);
        if (s1)
            finish_short(r, a, b, c1, v1);
        else
            long_row(r, a, b);
        if (has2) {
            if (s2)
                finish_short(r2, a2, b2, c2, v2);
            else
                long_row(r2, a2, b2);
        }
    }
    // one atomic per wave that found a problem
    int wbad = bad;
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) wbad |= __shfl_xor(wbad, o);
    if (wbad && lane == 0) atomicOr(flag, wbad);
}

// ---------------------------------------------------------------------------
// Omega on the device (the reference stream, scx-accel/src/pca/cpu.rs:1174-1178:
// ChaCha8Rng::seed_from_u64 + StandardNormal).
// k_chacha8_u64: u64 draw w = (u32[2w], u32[2w+1]) of the ChaCha8 keystream
// (8 rounds, 64-bit block counter in words 12-13, zero nonce), one 16-word
// block per thread.
// k_zig_outcome: for every draw position w, the outcome of a StandardNormal
// draw that starts there (rand_distr's 256-box ziggurat): the number of u64
// draws it consumes and its value. Flagged (bit 7) for the host to recompute
// with its libm: tail draws (their value goes through log), wedge tests whose
// exp comparison is within 4 ulp (the device exp may round differently), and
// draws that would leave the computed words. Everything else is IEEE
// multiplies and comparisons (no contraction), identical to the host.
// ---------------------------------------------------------------------------
struct ChaChaKey { uint32_t k[8]; };

__device__ __forceinline__ uint32_t rotl32(uint32_t v, int r) { return (v << r) | (v >> (32 - r)); }

__global__ void __launch_bounds__(256) k_chacha8_u64(ChaChaKey key, uint64_t n_words, uint64_t *__restrict__ out) {
    const uint64_t b = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (8 * b >= n_words) return;
    uint32_t s[16] = {0x61707865u, 0x3320646eu, 0x79622d32u, 0x6b206574u,
                      key.k[0], key.k[1], key.k[2], key.k[3], key.k[4], key.k[5], key.k[6], key.k[7],
                      static_cast<uint32_t>(b), static_cast<uint32_t>(b >> 32), 0u, 0u};
    uint32_t x[16];
#pragma unroll
    for (int i = 0; i < 16; ++i) x[i] = s[i];
#define SCX_QR(a, b2, c, d)                      \
    x[a] += x[b2]; x[d] = rotl32(x[d] ^ x[a], 16); \
    x[c] += x[d]; x[b2] = rotl32(x[b2] ^ x[c], 12); \
    x[a] += x[b2]; x[d] = rotl32(x[d] ^ x[a], 8);  \
    x[c] += x[d]; x[b2] = rotl32(x[b2] ^ x[c], 7);
#pragma unroll
    for (int dr = 0; dr < 4; ++dr) {
        SCX_QR(0, 4, 8, 12) SCX_QR(1, 5, 9, 13) SCX_QR(2, 6, 10, 14) SCX_QR(3, 7, 11, 15)
        SCX_QR(0, 5, 10, 15) SCX_QR(1, 6, 11, 12) SCX_QR(2, 7, 8, 13) SCX_QR(3, 4, 9, 14)
    }
#undef SCX_QR
#pragma unroll
    for (int i = 0; i < 8; ++i) {
        const uint64_t w = 8 * b + i;
        if (w < n_words)
            out[w] = static_cast<uint64_t>(x[2 * i] + s[2 * i]) |
                     (static_cast<uint64_t>(x[2 * i + 1] + s[2 * i + 1]) << 32);
    }
}

__global__ void __launch_bounds__(256) k_zig_outcome(const uint64_t *__restrict__ W, uint64_t n_words,
                                                     const double *__restrict__ X, const double *__restrict__ F,
                                                     uint8_t *__restrict__ c_out, double *__restrict__ v_out) {
#pragma clang fp contract(off)
    constexpr uint64_t kExp2 = static_cast<uint64_t>(1023 + 1) << 52;  // [2,4)
    const uint64_t stride = static_cast<uint64_t>(gridDim.x) * blockDim.x;
    for (uint64_t w = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x; w < n_words; w += stride) {
        uint32_t c = 0;
        bool flag = false;
        double val = 0.0;
        for (;;) {
            if (w + c >= n_words || c > 120) { flag = true; break; }
            const uint64_t bits = W[w + c];
            ++c;
            const unsigned i = static_cast<unsigned>(bits & 0xffu);
            const double u = __builtin_bit_cast(double, (bits >> 12) | kExp2) - 3.0;
            const double xv = u * X[i];
            if (fabs(xv) < X[i + 1]) { val = xv; break; }
            if (i == 0) { flag = true; break; }  // tail
            if (w + c >= n_words) { flag = true; break; }
            const double unif = static_cast<double>(W[w + c] >> 11) * 0x1p-53;
            ++c;
            const double lhs = F[i + 1] + (F[i] - F[i + 1]) * unif;
            const double pdf = exp(-xv * xv / 2.0);
            if (fabs(lhs - pdf) <= 4.0 * 2.220446049250313e-16 * pdf) { flag = true; break; }
            if (lhs < pdf) { val = xv; break; }
        }
        c_out[w] = flag ? static_cast<uint8_t>(0x80u) : static_cast<uint8_t>(c);
        v_out[w] = val;
    }
}

__global__ void k_omega_override(double *__restrict__ v, const uint32_t *__restrict__ pos,
                                 const double *__restrict__ val, int n) {
    const int i = static_cast<int>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < n) v[pos[i]] = val[i];
}

// Omega element (r, j) = draw r * k + j, as the f32 panel (KP columns, zero pad)
// and, when out64 is given, as f64 row-major rows x k.
template <int KP>
__global__ void __launch_bounds__(256) k_omega_panel(const double *__restrict__ v, const uint32_t *__restrict__ start,
                                                     int64_t rows, int k, float *__restrict__ panel,
                                                     double *__restrict__ out64) {
    const int64_t t = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (t >= rows * KP) return;
    const int64_t r = t / KP;
    const int j = static_cast<int>(t % KP);
    double x = 0.0;
    if (j < k) {
        x = v[start[r * k + j]];
        if (out64) out64[r * k + j] = x;
    }
    if (panel) panel[t] = static_cast<float>(x);
}

// ---------------------------------------------------------------------------
// Wide panels (k' > 128). A panel of kpw = nb * 128 columns is stored as nb
// blocks of 128 columns, block-major: block b of an n x kpw panel is the
// row-major n x 128 panel at offset b * n * 128, so every block is a panel the
// KP = 128 SpMM kernels take as it is. The library QR and SVD work on
// column-major n x k copies.
// ---------------------------------------------------------------------------
constexpr int kWideBlock = 128;

__host__ __device__ __forceinline__ int64_t wide_index(int64_t rows, int64_t r, int j) {
    return static_cast<int64_t>(j / kWideBlock) * rows * kWideBlock + r * kWideBlock + (j % kWideBlock);
}

// Omega element (r, j) = draw r * k + j into a block-major f32 panel (zero past k).
__global__ void __launch_bounds__(256) k_omega_blocks(const double *__restrict__ v, const uint32_t *__restrict__ start,
                                                      int64_t rows, int k, int kpw, float *__restrict__ panel) {
    const int64_t t = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (t >= rows * kpw) return;
    const int64_t r = t / kpw;
    const int j = static_cast<int>(t % kpw);
    panel[wide_index(rows, r, j)] = j < k ? static_cast<float>(v[start[r * k + j]]) : 0.f;
}

// dst (column-major rows x k, ld = rows) <- src (block-major), 64 x 64 tiles
// through LDS: rows read and columns written contiguously. Grid (rows / 64,
// k / 64), 256 threads.
template <typename TS, typename TD>
__global__ void __launch_bounds__(256) k_blocks_to_cols(const TS *__restrict__ src, int64_t rows, int k,
                                                        TD *__restrict__ dst) {
    __shared__ TD tile[64][65];
    const int64_t r0 = static_cast<int64_t>(blockIdx.x) * 64;
    const int j0 = static_cast<int>(blockIdx.y) * 64;
    const int tx = threadIdx.x & 63, ty = threadIdx.x >> 6;
    for (int i = ty; i < 64; i += 4) {
        const int64_t r = r0 + i;
        const int j = j0 + tx;
        tile[i][tx] = (r < rows && j < k) ? static_cast<TD>(src[wide_index(rows, r, j)]) : TD(0);
    }
    __syncthreads();
    for (int i = ty; i < 64; i += 4) {
        const int j = j0 + i;
        const int64_t r = r0 + tx;
        if (r < rows && j < k) dst[static_cast<int64_t>(j) * rows + r] = tile[tx][i];
    }
}

// dst (block-major, kpw columns, zero at columns >= k) <- src (column-major
// rows x k, ld = rows). Grid (rows / 64, kpw / 64), 256 threads.
template <typename T>
__global__ void __launch_bounds__(256) k_cols_to_blocks(const T *__restrict__ src, int64_t rows, int k, int kpw,
                                                        T *__restrict__ dst) {
    __shared__ T tile[64][65];
    const int64_t r0 = static_cast<int64_t>(blockIdx.x) * 64;
    const int j0 = static_cast<int>(blockIdx.y) * 64;
    const int tx = threadIdx.x & 63, ty = threadIdx.x >> 6;
    for (int i = ty; i < 64; i += 4) {
        const int j = j0 + i;
        const int64_t r = r0 + tx;
        tile[i][tx] = (r < rows && j < k) ? src[static_cast<int64_t>(j) * rows + r] : T(0);
    }
    __syncthreads();
    for (int i = ty; i < 64; i += 4) {
        const int64_t r = r0 + i;
        const int j = j0 + tx;
        if (r < rows && j < kpw) dst[wide_index(rows, r, j)] = tile[tx][i];
    }
}

// W (column-major k x nc, f32) = (V S)[:, :nc] from the SVD's V^T (column-major
// k x k) and the singular values S: the embedding weights of the wide panel.
__global__ void __launch_bounds__(256) k_wide_emb_weights(const double *__restrict__ Vt, const double *__restrict__ S,
                                                          int k, int nc, float *__restrict__ W) {
    const int64_t t = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (t >= static_cast<int64_t>(k) * nc) return;
    const int l = static_cast<int>(t % k), j = static_cast<int>(t / k);
    W[t] = static_cast<float>(Vt[j + static_cast<int64_t>(l) * k] * S[j]);
}

// status |= bit when a library routine reported a nonzero info word
__global__ void k_info_status(const int *__restrict__ info, int bit, int *__restrict__ status) {
    if (threadIdx.x == 0 && blockIdx.x == 0 && *info != 0) atomicOr(status, bit);
}

// columns [c0, c0 + ncols) of a row-major dst (ld columns) <- the first ncols
// columns of a KP-wide panel
template <int KP>
__global__ void __launch_bounds__(256) k_unpack_cols(const float *__restrict__ src, int64_t rows, int ncols,
                                                     float *__restrict__ dst, int ld, int c0) {
    const int64_t t = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (t >= rows * ncols) return;
    const int64_t r = t / ncols;
    const int j = static_cast<int>(t % ncols);
    dst[r * ld + c0 + j] = src[r * KP + j];
}

// p[0, n) = 0 (32-bit words): a fill kernel where a memset node is to be
// avoided (work a caller may capture into a HIP graph).
__global__ void __launch_bounds__(256) k_zero_words(uint32_t *__restrict__ p, int64_t n) {
    const int64_t stride = static_cast<int64_t>(gridDim.x) * blockDim.x;
    for (int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += stride) p[i] = 0u;
}

// bf16 (RNE) copy of an f32 array: the sketch form's X~ values.
// Packed bf16 tile entries (pack_bf16_entry) from the tile-CSC, for a matrix
// whose build did not write them: one wave per (tile, 64-gene group), whose
// runs are one contiguous range; lane j holds gene g0 + j's run start and
// each entry's gene comes from a 6-step search over them. Grid: 4 waves per
// block over n_tiles * ceil(n_vars / 64) groups.
__global__ void __launch_bounds__(256) k_pack_tile_entries(const int64_t *__restrict__ indptr,
                                                           const uint32_t *__restrict__ tptr,
                                                           const uint8_t *__restrict__ trow,
                                                           const float *__restrict__ tval, int64_t n_vars,
                                                           int n_tiles, uint32_t *__restrict__ tpk) {
    const int lane = threadIdx.x & 63;
    const int64_t ng = (n_vars + 63) / 64;
    const int64_t wid = static_cast<int64_t>(blockIdx.x) * (blockDim.x / 64) + (threadIdx.x >> 6);
    if (wid >= static_cast<int64_t>(n_tiles) * ng) return;  // wave-uniform
    const int64_t t = wid / ng, g0 = (wid % ng) * 64;
    const uint32_t *tp = tptr + t * (n_vars + 1);
    const int64_t gl = g0 + lane;
    const uint32_t st = tp[gl < n_vars ? gl : n_vars];
    const uint32_t end = tp[(g0 + 64) < n_vars ? g0 + 64 : n_vars];
    const uint32_t s0 = static_cast<uint32_t>(__builtin_amdgcn_readlane(static_cast<int>(st), 0));
    const int64_t tb = indptr[t * 256];
    for (uint32_t p0 = s0; p0 < end; p0 += 64) {
        const uint32_t p = p0 + static_cast<uint32_t>(lane);
        int lo = 0;
#pragma unroll
        for (int step = 32; step >= 1; step >>= 1) {
            const uint32_t v = static_cast<uint32_t>(__shfl(static_cast<int>(st), lo + step < 64 ? lo + step : 63));
            lo = (lo + step < 64 && v <= p) ? lo + step : lo;
        }
        if (p < end) tpk[tb + p] = pack_bf16_entry(tval[tb + p], trow[tb + p], static_cast<uint32_t>(g0 + lo));
    }
}

// lg[p] = the gene field of xpk[p]
__global__ void __launch_bounds__(256) k_unpack_row_genes(const uint32_t *__restrict__ xpk, int64_t nnz,
                                                          uint16_t *__restrict__ lg) {
    const int64_t p = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (p < nnz) lg[p] = static_cast<uint16_t>(xpk[p]);
}

// xpk[p] = pack_bf16_row_entry(xt[p], lg[p])
__global__ void __launch_bounds__(256) k_pack_row_entries(const float *__restrict__ xt,
                                                          const uint16_t *__restrict__ lg, int64_t nnz,
                                                          uint32_t *__restrict__ xpk) {
    const int64_t p = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (p < nnz) xpk[p] = pack_bf16_row_entry(xt[p], lg[p]);
}

__global__ void __launch_bounds__(256) k_to_bf16(const float *__restrict__ src, int64_t n,
                                                 uint16_t *__restrict__ out) {
    const int64_t stride = static_cast<int64_t>(gridDim.x) * blockDim.x;
    for (int64_t p = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x; p < n; p += stride)
        out[p] = f32_to_bf16_rne(src[p]);
}

// Column statistics partials -> totals, fixed order over row blocks.
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
                                                     const float *__restrict__ xt,
                                                     const double *__restrict__ mu, int64_t n_obs,
                                                     double *__restrict__ part) {
    __shared__ double red[256];
    const int lane = threadIdx.x & 63;
    double acc = 0.0;
    const int64_t wstride = static_cast<int64_t>(gridDim.x) * (blockDim.x / kWave);
    for (int64_t r = static_cast<int64_t>(blockIdx.x) * (blockDim.x / kWave) + threadIdx.x / kWave;
         r < n_obs; r += wstride) {
        const int64_t a = indptr[r], b = indptr[r + 1];
        for (int64_t p = a + lane; p < b; p += kWave) {
            const double d = static_cast<double>(xt[p]) - mu[idx[p]];
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

// (sum x, sum x^2) over x[0..n), f64, deterministic: block b of the first
// launch (gridDim.x = nparts) sums its contiguous chunk (thread partials over
// a fixed stride, then a fixed-order tree) into part[b]; the second launch
// (one block, x = part, n = nparts, part = null) reduces the parts, read as
// (sum, sum of squares) pairs, into out.
__global__ void __launch_bounds__(256) k_sum_sq_f64(const double *__restrict__ x, int64_t n,
                                                    double *__restrict__ part, double *__restrict__ out) {
    __shared__ double rs[256], rq[256];
    double a = 0.0, q = 0.0;
    if (part) {
        const int64_t chunk = (n + gridDim.x - 1) / gridDim.x;
        const int64_t b0 = static_cast<int64_t>(blockIdx.x) * chunk;
        const int64_t b1 = b0 + chunk < n ? b0 + chunk : n;
        for (int64_t i = b0 + threadIdx.x; i < b1; i += 256) {
            const double v = x[i];
            a += v;
            q += v * v;
        }
    } else {
        for (int64_t i = threadIdx.x; i < n; i += 256) {
            a += x[2 * i];
            q += x[2 * i + 1];
        }
    }
    rs[threadIdx.x] = a;
    rq[threadIdx.x] = q;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if (static_cast<int>(threadIdx.x) < w) {
            rs[threadIdx.x] += rs[threadIdx.x + w];
            rq[threadIdx.x] += rq[threadIdx.x + w];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        double *o = part ? part + 2 * blockIdx.x : out;
        o[0] = rs[0];
        o[1] = rq[0];
    }
}

// part[blk][j] = sum over the block's rows r of w[r] * A[r][j] (w == null:
// 1), f64, for a tall f32 panel A (nrows x KP): 16-byte loads (KP / 4 lanes per
// row), four rows per lane in flight, the row lanes reduced in a fixed order.
template <int KP>
__global__ void __launch_bounds__(256) k_wcolsum4(const float *__restrict__ A, const double *__restrict__ w,
                                                  int64_t nrows, int64_t rows_per_blk, double *__restrict__ part) {
    constexpr int CG = KP / 4, RL = 256 / CG;
    __shared__ double red[RL][KP];
    const int cg = threadIdx.x % CG, rl = threadIdx.x / CG;
    const int64_t r0 = static_cast<int64_t>(blockIdx.x) * rows_per_blk;
    const int64_t r1 = r0 + rows_per_blk < nrows ? r0 + rows_per_blk : nrows;
    double acc[4] = {0.0, 0.0, 0.0, 0.0};
    int64_t r = r0 + rl;
    for (; r + 3 * RL < r1; r += 4 * RL) {
        float4 v[4];
        double ww[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            v[u] = *reinterpret_cast<const float4 *>(A + (r + u * RL) * KP + 4 * cg);
            ww[u] = w ? w[r + u * RL] : 1.0;
        }
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            acc[0] += ww[u] * static_cast<double>(v[u].x);
            acc[1] += ww[u] * static_cast<double>(v[u].y);
            acc[2] += ww[u] * static_cast<double>(v[u].z);
            acc[3] += ww[u] * static_cast<double>(v[u].w);
        }
    }
    for (; r < r1; r += RL) {
        const float4 v = *reinterpret_cast<const float4 *>(A + r * KP + 4 * cg);
        const double ww = w ? w[r] : 1.0;
        acc[0] += ww * static_cast<double>(v.x);
        acc[1] += ww * static_cast<double>(v.y);
        acc[2] += ww * static_cast<double>(v.z);
        acc[3] += ww * static_cast<double>(v.w);
    }
#pragma unroll
    for (int e = 0; e < 4; ++e) red[rl][4 * cg + e] = acc[e];
    __syncthreads();
    if (threadIdx.x < KP) {
        double t = 0.0;
        for (int i = 0; i < RL; ++i) t += red[i][threadIdx.x];
        part[static_cast<int64_t>(blockIdx.x) * KP + threadIdx.x] = t;
    }
}

// Column j of the partials summed by block j: thread t takes partials t,
// t+256, ...; then a fixed-order tree over the 256 thread sums.
template <int KP>
__global__ void __launch_bounds__(256) k_colsum_reduce(const double *__restrict__ part, int nblk,
                                                       double *__restrict__ out64, float *__restrict__ out32) {
    __shared__ double red[256];
    const int j = blockIdx.x;
    double acc = 0.0;
    for (int b = threadIdx.x; b < nblk; b += 256) acc += part[static_cast<int64_t>(b) * KP + j];
    red[threadIdx.x] = acc;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if (static_cast<int>(threadIdx.x) < w) red[threadIdx.x] += red[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        if (out64) out64[j] = red[0];
        if (out32) out32[j] = static_cast<float>(red[0]);
    }
}

// ---------------------------------------------------------------------------
// Forward SpMM: Y[r][:] = sum_{nz in row r} x~ * M[col][:] - mc   (cpu.rs:1471-1523)
// One wavefront per row; a batch of 64 nonzeros is transformed lane-parallel
// and staged in LDS; the four 16-lane groups then take nonzeros 4t+g and
// gather M rows (L2-resident panel) as float4; the groups' partial rows are
// summed with cross-lane shuffles.
// ---------------------------------------------------------------------------

// BF16: x~ rounded to bf16 (the BF16 sketch form on the general layout)
template <int KP, bool OFFSET = false, bool BF16 = false>  // OFFSET: + rowb[r] cs[j] (PFlog)
__global__ void __launch_bounds__(256) k_spmm_fwd(const int64_t *__restrict__ indptr,
                                                  const uint32_t *__restrict__ idx,
                                                  const float *__restrict__ xt,
                                                  const float *__restrict__ M, const float *__restrict__ mc,
                                                  float *__restrict__ Y, int64_t n_obs,
                                                  const double *__restrict__ rowb, const float *__restrict__ cs) {
    constexpr int E = KP / 16;
    __shared__ uint2 stage[4][kWave];
    const int wave = threadIdx.x / kWave, lane = threadIdx.x & 63, g = lane >> 4, l = lane & 15;
    const int64_t wstride = static_cast<int64_t>(gridDim.x) * 4;
    for (int64_t r = static_cast<int64_t>(blockIdx.x) * 4 + wave; r < n_obs; r += wstride) {
        const int64_t a = indptr[r], b = indptr[r + 1];
        float acc[E];
#pragma unroll
        for (int e = 0; e < E; ++e) acc[e] = 0.f;
        for (int64_t base = a; base < b; base += kWave) {
            const int cnt = (b - base) < kWave ? static_cast<int>(b - base) : kWave;
            uint32_t c = 0u;
            float x = 0.f;
            if (lane < cnt) {
                c = idx[base + lane];
                x = xt[base + lane];
            }
            stage[wave][lane] = make_uint2(c, __float_as_uint(x));
            wave_lds_sync();
            const int steps = (cnt + 3) >> 2;
            for (int t = 0; t < steps; ++t) {
                const uint2 ent = stage[wave][4 * t + g];
                const float xv = BF16 ? bf16_round(__uint_as_float(ent.y)) : __uint_as_float(ent.y);
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
            const float br = OFFSET ? static_cast<float>(rowb[r]) : 0.f;
#pragma unroll
            for (int e = 0; e < E / 4; ++e) {
                float4 o;
                o.x = acc[4 * e + 0] - (mc ? mc[l * E + 4 * e + 0] : 0.f);
                o.y = acc[4 * e + 1] - (mc ? mc[l * E + 4 * e + 1] : 0.f);
                o.z = acc[4 * e + 2] - (mc ? mc[l * E + 4 * e + 2] : 0.f);
                o.w = acc[4 * e + 3] - (mc ? mc[l * E + 4 * e + 3] : 0.f);
                if constexpr (OFFSET) {
                    o.x += br * cs[l * E + 4 * e + 0];
                    o.y += br * cs[l * E + 4 * e + 1];
                    o.z += br * cs[l * E + 4 * e + 2];
                    o.w += br * cs[l * E + 4 * e + 3];
                }
                yr[e] = o;
            }
        }
    }
}

// B[v][j] = sum_rb part[rb][v][j] - mu[v] * sumq[j]  (f64; cpu.rs:1253-1267),
// + bq[j] = (b^T Q)_j for a PFlog matrix (cpu.rs:653-687)
template <int KP>
__global__ void __launch_bounds__(256) k_spmm_t_reduce(const float *__restrict__ part, int nrb,
                                                       int64_t n_vars, const double *__restrict__ mu,
                                                       const double *__restrict__ sumq,
                                                       double *__restrict__ B, float *__restrict__ B32,
                                                       const double *__restrict__ bq) {
    const int64_t t = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (t >= n_vars * KP) return;
    const int64_t v = t / KP;
    const int j = static_cast<int>(t % KP);
    const int64_t stride = n_vars * KP;
    double s = 0.0;
    int rb = 0;
    for (; rb + 8 <= nrb; rb += 8) {  // eight loads in flight, summed in partial order
        float p[8];
#pragma unroll
        for (int u = 0; u < 8; ++u) p[u] = part[static_cast<int64_t>(rb + u) * stride + t];
#pragma unroll
        for (int u = 0; u < 8; ++u) s += static_cast<double>(p[u]);
    }
    for (; rb < nrb; ++rb) s += static_cast<double>(part[static_cast<int64_t>(rb) * stride + t]);
    if (mu) s -= mu[v] * sumq[j];
    if (bq) s += bq[j];
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

// G[i][j] (full symmetric) from the upper-tile partials. Block b reduces the
// 64 entries [64b, 64b+64): 4 thread groups take every 4th partial, then the 4
// group sums are added in a fixed order (deterministic).
template <int KP>
__global__ void __launch_bounds__(256) k_gram_reduce(const double *__restrict__ part, int nparts,
                                                     double *__restrict__ G) {
    __shared__ double red[4][64];
    const int t = blockIdx.x * 64 + (threadIdx.x & 63);
    const int grp = threadIdx.x >> 6;
    int i = t / KP, j = t % KP;
    if (i / 32 > j / 32) { const int s = i; i = j; j = s; }
    double acc = 0.0;
    if (t < KP * KP)
        for (int p = grp; p < nparts; p += 4) acc += part[static_cast<int64_t>(p) * KP * KP + i * KP + j];
    red[grp][threadIdx.x & 63] = acc;
    __syncthreads();
    if (grp == 0 && t < KP * KP) G[t] = ((red[0][threadIdx.x] + red[1][threadIdx.x]) + red[2][threadIdx.x]) + red[3][threadIdx.x];
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


// out = A W for a small f64 A (rows x KP), W f64 KP x KP: W and the block's
// 256 / KP rows of A staged in LDS, one output element per thread (four
// partial sums over the inner index, added in a fixed order).
template <int KP>
__global__ void __launch_bounds__(256) k_apply64(const double *__restrict__ A, const double *__restrict__ W,
                                                 double *__restrict__ out, float *__restrict__ out32,
                                                 int64_t nrows) {
    constexpr int RB = 256 / KP;  // rows per block
    __shared__ double ws[KP * KP];
    __shared__ double as[RB * KP];
    const int64_t r0 = static_cast<int64_t>(blockIdx.x) * RB;
    for (int i = threadIdx.x; i < KP * KP; i += blockDim.x) ws[i] = W[i];
    for (int i = threadIdx.x; i < RB * KP; i += blockDim.x)
        as[i] = r0 + i / KP < nrows ? A[r0 * KP + i] : 0.0;
    __syncthreads();
    const int rl = threadIdx.x / KP, j = threadIdx.x % KP;
    if (r0 + rl >= nrows) return;
    double s0 = 0.0, s1 = 0.0, s2 = 0.0, s3 = 0.0;
#pragma unroll 4
    for (int i = 0; i < KP; i += 4) {
        s0 += as[rl * KP + i] * ws[i * KP + j];
        s1 += as[rl * KP + i + 1] * ws[(i + 1) * KP + j];
        s2 += as[rl * KP + i + 2] * ws[(i + 2) * KP + j];
        s3 += as[rl * KP + i + 3] * ws[(i + 3) * KP + j];
    }
    const double sm = (s0 + s1) + (s2 + s3);
    const int64_t t = (r0 + rl) * KP + j;
    if (out) out[t] = sm;
    if (out32) out32[t] = static_cast<float>(sm);
}

// C = A B for KP x KP f64 matrices; block b computes rows 16b .. 16b + 15
// (grid KP / 16): those rows of A staged in LDS, B staged whole for KP = 64
// (32 KB) and read from global memory for KP = 128; each product summed over
// the inner index as two interleaved partial sums (even, odd), then added.
template <int KP>
__global__ void __launch_bounds__(256) k_small_gemm64(const double *__restrict__ A, const double *__restrict__ B,
                                                      double *__restrict__ C) {
    static_assert(KP == 64 || KP == 128, "panel widths");
    constexpr int RT = 16;  // rows of C per block
    __shared__ double bs[KP == 64 ? KP * KP : 1];
    __shared__ double at[RT * KP];
    const int i0 = static_cast<int>(blockIdx.x) * RT;
    if constexpr (KP == 64)
        for (int i = threadIdx.x; i < KP * KP; i += blockDim.x) bs[i] = B[i];
    for (int i = threadIdx.x; i < RT * KP; i += blockDim.x) at[i] = A[i0 * KP + i];
    __syncthreads();
    for (int t = threadIdx.x; t < RT * KP; t += blockDim.x) {
        const int i = t / KP, j = t % KP;
        double s0 = 0.0, s1 = 0.0;
        for (int q = 0; q < KP; q += 2) {
            const double b0 = KP == 64 ? bs[q * KP + j] : B[q * KP + j];
            const double b1 = KP == 64 ? bs[(q + 1) * KP + j] : B[(q + 1) * KP + j];
            s0 += at[i * KP + q] * b0;
            s1 += at[i * KP + q + 1] * b1;
        }
        C[(i0 + i) * KP + j] = s0 + s1;
    }
}

// Embedding weights W_emb = P[:, :nc] Sigma as f32 KP x KP (zero padded),
// from P = Rinv V (k_small_gemm64), Rinv the last CholeskyQR pass's inverse
// of the final basis: the embeddings Q W_emb of Q = Q1 Rinv come from Q1.
template <int KP>
__global__ void __launch_bounds__(256) k_emb_weights(const double *__restrict__ P, const double *__restrict__ sigma,
                                                     int nc, float *__restrict__ W) {
    for (int t = threadIdx.x; t < KP * KP; t += blockDim.x) {
        const int j = t % KP;
        W[t] = j < nc ? static_cast<float>(P[t] * sigma[j]) : 0.f;
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

// sig (kp), the column sums and sums of squares (nv each) and the status word
// into one f64 buffer, so the end of a PCA call copies back once
__global__ void __launch_bounds__(256) k_pack_results(const double *__restrict__ sig, int kp,
                                                      const double *__restrict__ sum, const double *__restrict__ sq,
                                                      int64_t nv, const int *__restrict__ status,
                                                      double *__restrict__ out) {
    const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= kp + 2 * nv + 1) return;
    double v;
    if (i < kp)
        v = sig[i];
    else if (i < kp + nv)
        v = sum[i - kp];
    else if (i < kp + 2 * nv)
        v = sq[i - kp - nv];
    else
        v = static_cast<double>(*status);
    out[i] = v;
}

__global__ void k_f32_to_f64(const float *__restrict__ s, double *__restrict__ d, int64_t n) {
    const int64_t t = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (t < n) d[t] = static_cast<double>(s[t]);
}

__global__ void k_f64_to_f32(const double *__restrict__ s, float *__restrict__ d, int64_t n) {
    const int64_t t = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (t < n) d[t] = static_cast<float>(s[t]);
}


// ===========================================================================
// Second-generation SpMM kernels.
// ===========================================================================

// Lane t of each 16-lane row broadcast to the whole row (DPP row_newbcast, gfx90a+).

// Gather step primitives. Each 16-lane group holds a batch of its own
// entries, entry T in lane T (an LDS byte offset of the panel row, and x~).
// Step T of a batch: one v_add_u32_dpp (row_newbcast:T) forms every lane's
// LDS address (lane T's row offset + the lane's column offset), ds_read_b128
// loads E/4 x 4 floats of the panel row, and each v_fmac_f32_dpp takes the
// broadcast x~ of lane T as its DPP source, so a step costs E/4 LDS reads,
// E FMAs and one add for the four groups' four entries. The s_nop 1 in each
// block covers the VALU-write -> DPP-read hazard (two wait states) for
// operands the compiler may have just moved.
template <int T>
__device__ __forceinline__ uint32_t bcast_add(uint32_t ofs, uint32_t lane_ofs) {
    uint32_t a;
    asm("s_nop 1\n\tv_add_u32_dpp %0, %1, %2 row_newbcast:%3 row_mask:0xf bank_mask:0xf bound_ctrl:1"
        : "=v"(a)
        : "v"(ofs), "v"(lane_ofs), "i"(T));
    return a;
}

template <int T>
__device__ __forceinline__ void bcast_fma4(float &a0, float &a1, float &a2, float &a3, float x, const float4 &m) {
    asm("s_nop 1\n\t"
        "v_fmac_f32_dpp %0, %4, %5 row_newbcast:%9 row_mask:0xf bank_mask:0xf bound_ctrl:1\n\t"
        "v_fmac_f32_dpp %1, %4, %6 row_newbcast:%9 row_mask:0xf bank_mask:0xf bound_ctrl:1\n\t"
        "v_fmac_f32_dpp %2, %4, %7 row_newbcast:%9 row_mask:0xf bank_mask:0xf bound_ctrl:1\n\t"
        "v_fmac_f32_dpp %3, %4, %8 row_newbcast:%9 row_mask:0xf bank_mask:0xf bound_ctrl:1"
        : "+v"(a0), "+v"(a1), "+v"(a2), "+v"(a3)
        : "v"(x), "v"(m.x), "v"(m.y), "v"(m.z), "v"(m.w), "i"(T));
}

template <int T, int E>
__device__ __forceinline__ void gather_load(float4 (&m)[E / 4], uint32_t ofs, const char *lds, uint32_t lane_ofs) {
    const uint32_t a = bcast_add<T>(ofs, lane_ofs);
#pragma unroll
    for (int e = 0; e < E / 4; ++e) m[e] = *reinterpret_cast<const float4 *>(lds + a + 16 * e);
}

template <int T, int E>
__device__ __forceinline__ void gather_fma(float (&acc)[E], const float4 (&m)[E / 4], float x) {
#pragma unroll
    for (int e = 0; e < E / 4; ++e)
        bcast_fma4<T>(acc[4 * e + 0], acc[4 * e + 1], acc[4 * e + 2], acc[4 * e + 3], x, m[e]);
}

constexpr int kTileRows = 256;

constexpr int kTcscGenes = 4096;   // genes per mask pass: 8 words each (256-row tile)
constexpr int kTcscStage = 12288;  // staged output entries (5 B each: 60 KB)

// Deterministic build. Per pass over a gene range, each gene gets a 256-bit
// row-membership mask; wave w owns rows [32w, 32w+32) of the tile, i.e. mask
// word w of every gene, so the masks are built without atomics. A nonzero
// (r, c) lands at start[c] + popcount(mask[c] below bit r): runs come out in
// row order whatever the schedule. When the pass's output fits kTcscStage
// entries it is assembled in LDS and written out contiguously.
__global__ void __launch_bounds__(512) k_build_tcsc(const int64_t *__restrict__ indptr,
                                                    const uint32_t *__restrict__ idx,
                                                    const float *__restrict__ val, int64_t n_obs,
                                                    int n_vars, uint32_t *__restrict__ tptr,
                                                    uint8_t *__restrict__ trow, float *__restrict__ tval,
                                                    const int *__restrict__ flag) {
    constexpr int GP = (kTcscGenes < 2048) ? kTcscGenes : 2048;  // genes per pass with staging room
    if (*flag) return;  // the matrix failed validation: nothing here may trust its indices
    __shared__ uint32_t mask[GP * 8];
    __shared__ uint32_t start[GP];
    __shared__ float sval[kTcscStage];
    __shared__ uint8_t srow[kTcscStage];
    __shared__ uint32_t wsum[80];
    __shared__ uint32_t carry;
    const int t = blockIdx.x;
    const int64_t r0 = static_cast<int64_t>(t) * kTileRows;
    const int64_t r1 = r0 + kTileRows < n_obs ? r0 + kTileRows : n_obs;
    const int64_t base = indptr[r0];
    const int wave = threadIdx.x / kWave, lane = threadIdx.x & 63, nw = blockDim.x / kWave;  // nw = 8
    uint32_t *tp = tptr + static_cast<int64_t>(t) * (n_vars + 1);
    if (threadIdx.x == 0) carry = 0u;
    for (int G0 = 0; G0 < n_vars; G0 += GP) {
        const int G1 = G0 + GP < n_vars ? G0 + GP : n_vars;
        const int ng = G1 - G0;
        for (int i = threadIdx.x; i < ng * 8; i += blockDim.x) mask[i] = 0u;
        __syncthreads();
        // wave w: rows 32w .. 32w+31 of the tile (mask word w), in order
        for (int rr = 32 * wave; rr < 32 * wave + 32; ++rr) {
            const int64_t r = r0 + rr;
            if (r >= r1) break;
            const uint32_t bit = 1u << (rr & 31);
            for (int64_t p = indptr[r] + lane; p < indptr[r + 1]; p += kWave) {
                const int c = static_cast<int>(idx[p]) - G0;
                if (c >= 0 && c < ng) mask[c * 8 + wave] |= bit;
            }
            wave_lds_sync();
        }
        __syncthreads();
        // exclusive scan of per-gene counts (contiguous segment per thread)
        const int per = (ng + blockDim.x - 1) / blockDim.x;
        const int c0 = threadIdx.x * per, c1 = c0 + per < ng ? c0 + per : ng;
        uint32_t local = 0;
        for (int c = c0; c < c1; ++c)
            for (int w = 0; w < 8; ++w) local += __popc(mask[c * 8 + w]);
        uint32_t incl = local;
        for (int o = 1; o < kWave; o <<= 1) {
            const uint32_t v = __shfl_up(incl, o);
            if (lane >= o) incl += v;
        }
        if (lane == kWave - 1) wsum[wave] = incl;
        __syncthreads();
        if (wave == 0) {
            const uint32_t wt = lane < nw ? wsum[lane] : 0u;
            uint32_t wincl = wt;
            for (int o = 1; o < kWave; o <<= 1) {
                const uint32_t v = __shfl_up(wincl, o);
                if (lane >= o) wincl += v;
            }
            if (lane < nw) wsum[32 + lane] = wincl - wt;
            if (lane == nw - 1) wsum[64] = wincl;
        }
        __syncthreads();
        const uint32_t pass_base = carry;
        const uint32_t pass_total = wsum[64];
        uint32_t off = wsum[32 + wave] + (incl - local);  // relative to the pass
        for (int c = c0; c < c1; ++c) {
            start[c] = off;
            tp[G0 + c] = pass_base + off;
            for (int w = 0; w < 8; ++w) off += __popc(mask[c * 8 + w]);
        }
        __syncthreads();
        const bool staged = pass_total <= static_cast<uint32_t>(kTcscStage);
        for (int rr = wave; rr < kTileRows; rr += nw) {
            const int64_t r = r0 + rr;
            if (r >= r1) break;
            const int word = rr >> 5;
            const uint32_t below = (rr & 31) ? ((1u << (rr & 31)) - 1u) : 0u;
            for (int64_t p = indptr[r] + lane; p < indptr[r + 1]; p += kWave) {
                const int c = static_cast<int>(idx[p]) - G0;
                if (c >= 0 && c < ng) {
                    uint32_t rank = __popc(mask[c * 8 + word] & below);
                    for (int w = 0; w < word; ++w) rank += __popc(mask[c * 8 + w]);
                    const uint32_t pos = start[c] + rank;
                    if (staged) {
                        srow[pos] = static_cast<uint8_t>(rr);
                        sval[pos] = val[p];
                    } else {
                        trow[base + pass_base + pos] = static_cast<uint8_t>(rr);
                        tval[base + pass_base + pos] = val[p];
                    }
                }
            }
        }
        __syncthreads();
        if (staged) {
            for (uint32_t i = threadIdx.x; i < pass_total; i += blockDim.x) {
                trow[base + pass_base + i] = srow[i];
                tval[base + pass_base + i] = sval[i];
            }
        }
        if (threadIdx.x == 0) carry = pass_base + pass_total;
        __syncthreads();
    }
    if (threadIdx.x == 0) tp[n_vars] = carry;
}

// Tile-CSC build from the split offsets (n_vars <= 8192), one block of 8
// waves per (tile, 256-gene chunk), XCD-aware (workgroups are dealt to the 8
// XCDs round-robin; block b runs on XCD b % 8 and takes tile
// 8 (b / 8 / nck) + b % 8, chunk (b / 8) % nck, so all chunks of a tile, whose
// row segments share cache lines, go through one L2; blocks past the last
// tile exit). Wave w owns the tile's rows 32w .. 32w+31 and flattens
// their segments of the chunk (CSR positions indptr[r] + [spl[r][2c],
// spl[r][2c+2]) for 128-gene split offsets); entry f = 64b + lane of batch b
// finds its row by a lane-parallel binary search over the 32 segment starts,
// and the wave keeps up to kB4Batches batches (gene, X~, row) in
// registers, so each entry is read from memory once. Pass 1 sets bit rl of
// mask[w][gene] (rows within a wave are distinct bits; a gene occurs at most
// once per row, which the matrix check enforces); the block then scans the
// per-gene counts into run starts and per-wave offsets, and pass 2 places each
// entry at woff[w][gene] + popcount(mask bits of earlier rows of the wave):
// runs in row order whatever the schedule (deterministic). The chunk is
// assembled in LDS and written contiguously when it has at most kB4Stage
// entries, directly otherwise. Longer waves process their remaining batches
// from memory in both passes. With psum, the block also writes the tile's
// per-gene sum and sum of squares of X~ (f64, run order).
// PK: also writes the packed bf16 entries (tpk), which stage a gene byte per
// entry: a 6,400-entry stage then, 7,680 without (16-bit wave offsets; either
// way 51 KB of LDS, three blocks per CU; about 1 % of the 1M x 2k matrix's
// (tile, chunk) pairs hold more than 6,400 entries and are written directly).
constexpr int kB4Genes = 256;
constexpr int kB4Batches = 12;
template <bool PK>
__global__ void __launch_bounds__(512) k_build_tcsc4(const int64_t *__restrict__ indptr,
                                                     const uint32_t *__restrict__ idx,
                                                     const float *__restrict__ xt,
                                                     const uint32_t *__restrict__ spl, int nsplit, int nck,
                                                     int64_t n_obs, int64_t n_vars, uint32_t *__restrict__ tptr,
                                                     uint8_t *__restrict__ trow, float *__restrict__ tval,
                                                     double *__restrict__ psum, double *__restrict__ psq,
                                                     const int *__restrict__ flag, uint32_t *__restrict__ tpk) {
    constexpr int NG = kB4Genes, NW = 8, U = kB4Batches;
    constexpr int kB4Stage = PK ? 6400 : 7680;
    if (*flag) return;  // the matrix failed validation: nothing here may trust its indices
    constexpr int SPC = NG / 128;  // 128-gene split offsets per chunk
    __shared__ uint32_t mask[NW][NG];
    __shared__ uint16_t woff[NW][NG];  // placement offsets < entries of the chunk <= 256 * 256
    __shared__ uint32_t start[NG + 1];
    __shared__ float sval[kB4Stage];
    __shared__ uint8_t srow[kB4Stage];
    __shared__ uint8_t sgene[PK ? kB4Stage : 1];  // gene in the chunk, for the packed bf16 entries
    __shared__ uint32_t wlo[NW], gsum[4];
    const int slot = static_cast<int>(blockIdx.x >> 3), xcd = static_cast<int>(blockIdx.x & 7);
    const int t = (slot / nck) * 8 + xcd, ck = slot % nck;
    const int64_t r0 = static_cast<int64_t>(t) * kTileRows;
    if (r0 >= n_obs) return;  // padding blocks of the last XCD round
    const int nr = (n_obs - r0) < kTileRows ? static_cast<int>(n_obs - r0) : kTileRows;
    const int tid = threadIdx.x, lane = tid & 63, w = __builtin_amdgcn_readfirstlane(tid >> 6);
    const uint32_t g0 = static_cast<uint32_t>(ck) * NG;
    const int ng = (n_vars - g0) < NG ? static_cast<int>(n_vars - g0) : NG;
    const int cs = ck * SPC, ce = (cs + SPC) < nsplit ? cs + SPC : nsplit;
    const int64_t base = indptr[r0];
    const int64_t p_last = indptr[n_obs] > 0 ? indptr[n_obs] - 1 : 0;
    // lanes 0..31: this chunk's segment of row 32w + lane
    uint32_t len = 0u, slo = 0u;
    int64_t gp = 0;
    if (lane < 32 && 32 * w + lane < nr) {
        const uint32_t *sr = spl + (r0 + 32 * w + lane) * (nsplit + 1);
        slo = sr[cs];
        len = sr[ce] - slo;
        gp = indptr[r0 + 32 * w + lane] + slo;
    }
    uint32_t incl = len;
#pragma unroll
    for (int o = 1; o < 32; o <<= 1) {
        const uint32_t v = __shfl_up(incl, o);
        if (lane >= o) incl += v;
    }
    const uint32_t fst = incl - len;  // flattened start of the row's segment
    const uint32_t wtotal = static_cast<uint32_t>(__builtin_amdgcn_readlane(static_cast<int>(incl), 31));
    const uint32_t wslo = static_cast<uint32_t>(wave_sum(static_cast<int>(slo)));
    for (int i = tid; i < NW * NG; i += blockDim.x) (&mask[0][0])[i] = 0u;
    if (lane == 0) wlo[w] = wslo;
    const int nb = static_cast<int>((wtotal + 63) / 64);
    // batch b: the lane's row (-1: none) and CSR position. Entry f = 64b + lane
    // belongs to the largest row j with fst[j] <= f (a 5-step binary search
    // over the 32 segment starts, read across lanes by ds_bpermute; an empty
    // row is never the largest such j unless it is the last, which the bound
    // check then rejects), when f < fst[j] + len[j].
    auto lane_u32 = [](int src_lane, uint32_t v) {
        return static_cast<uint32_t>(__builtin_amdgcn_ds_bpermute(src_lane * 4, static_cast<int>(v)));
    };
    const uint32_t gp_lo = static_cast<uint32_t>(static_cast<uint64_t>(gp));
    const uint32_t gp_hi = static_cast<uint32_t>(static_cast<uint64_t>(gp) >> 32);
    auto locate = [&](int b, int &row, int64_t &p) {
        const uint32_t f = 64u * static_cast<uint32_t>(b) + static_cast<uint32_t>(lane);
        int lo = 0;
        uint32_t fs = 0u;  // fst[0] = 0 <= f
#pragma unroll
        for (int step = 16; step > 0; step >>= 1) {
            const uint32_t v = lane_u32(lo + step, fst);
            const bool le = v <= f;
            lo = le ? lo + step : lo;
            fs = le ? v : fs;
        }
        const uint32_t ln = lane_u32(lo, len);
        const uint64_t g = (static_cast<uint64_t>(lane_u32(lo, gp_hi)) << 32) | lane_u32(lo, gp_lo);
        const bool in = f < fs + ln;
        row = in ? lo : -1;
        p = in ? static_cast<int64_t>(g) + (f - fs) : base;
        p = p < p_last ? p : p_last;
    };
    int rowv[U];
    uint32_t gv[U];
    float xv[U];
#pragma unroll
    for (int b = 0; b < U; ++b) {
        rowv[b] = -1;
        gv[b] = g0;
        xv[b] = 0.f;
        if (b < nb) {
            int64_t p;
            locate(b, rowv[b], p);
            gv[b] = idx[p];
            xv[b] = xt[p];
        }
    }
    __syncthreads();  // masks cleared
    // pass 1: row masks per gene
#pragma unroll
    for (int b = 0; b < U; ++b)
        if (rowv[b] >= 0) atomicOr(&mask[w][gv[b] - g0], 1u << rowv[b]);
    for (int b = U; b < nb; ++b) {
        int row;
        int64_t p;
        locate(b, row, p);
        const uint32_t g = idx[p];
        if (row >= 0) atomicOr(&mask[w][g - g0], 1u << row);
    }
    __syncthreads();
    // per-gene counts: run starts (block scan over genes) and per-wave offsets
    uint32_t tot = 0u;
    if (tid < NG) {
#pragma unroll
        for (int ww = 0; ww < NW; ++ww) {
            woff[ww][tid] = static_cast<uint16_t>(tot);
            tot += __popc(mask[ww][tid]);
        }
    }
    uint32_t ci = tot;
#pragma unroll
    for (int o = 1; o < kWave; o <<= 1) {
        const uint32_t v = __shfl_up(ci, o);
        if (lane >= o) ci += v;
    }
    if (lane == kWave - 1 && w < 4) gsum[w] = ci;
    __syncthreads();
    uint32_t cbase = 0u;
#pragma unroll
    for (int ww = 0; ww < NW; ++ww) cbase += wlo[ww];
    const uint32_t ctotal = gsum[0] + gsum[1] + gsum[2] + gsum[3];
    if (tid < NG) {
        uint32_t st = ci - tot;
        for (int ww = 0; ww < w; ++ww) st += gsum[ww];
        start[tid] = st;
#pragma unroll
        for (int ww = 0; ww < NW; ++ww) woff[ww][tid] = static_cast<uint16_t>(woff[ww][tid] + st);
        if (tid < ng) tptr[static_cast<int64_t>(t) * (n_vars + 1) + g0 + tid] = cbase + st;
    }
    if (tid == 0) {
        start[NG] = ctotal;
        if (ck == nck - 1) tptr[static_cast<int64_t>(t) * (n_vars + 1) + n_vars] = static_cast<uint32_t>(indptr[r0 + nr] - base);
    }
    __syncthreads();
    // pass 2: placement
    const bool staged = ctotal <= static_cast<uint32_t>(kB4Stage);
    auto place = [&](int row, uint32_t g, float x) {
        const uint32_t gl = g - g0;
        const uint32_t below = row ? (mask[w][gl] & ((1u << row) - 1u)) : 0u;
        const uint32_t pos = static_cast<uint32_t>(woff[w][gl]) + __popc(below);
        const uint8_t rt = static_cast<uint8_t>(32 * w + row);
        if (staged) {
            srow[pos] = rt;
            sval[pos] = x;
            if constexpr (PK) sgene[pos] = static_cast<uint8_t>(gl);
        } else {
            trow[base + cbase + pos] = rt;
            tval[base + cbase + pos] = x;
            if constexpr (PK) tpk[base + cbase + pos] = pack_bf16_entry(x, rt, gl);
        }
    };
#pragma unroll
    for (int b = 0; b < U; ++b)
        if (rowv[b] >= 0) place(rowv[b], gv[b], xv[b]);
    for (int b = U; b < nb; ++b) {
        int row;
        int64_t p;
        locate(b, row, p);
        const uint32_t g = idx[p];
        const float x = xt[p];
        if (row >= 0) place(row, g, x);
    }
    __syncthreads();
    if (staged)
        for (uint32_t q = tid; q < ctotal; q += blockDim.x) {
            trow[base + cbase + q] = srow[q];
            tval[base + cbase + q] = sval[q];
            if constexpr (PK) tpk[base + cbase + q] = pack_bf16_entry(sval[q], srow[q], sgene[q]);
        }
    if (psum && tid < ng) {
        const uint32_t a = start[tid], e = start[tid + 1];
        double sm = 0.0, sq = 0.0;
        uint32_t q = a;
        for (; q + 4 <= e; q += 4) {  // four loads in flight, the sums in run order
            double v[4];
#pragma unroll
            for (int u = 0; u < 4; ++u) v[u] = static_cast<double>(staged ? sval[q + u] : tval[base + cbase + q + u]);
#pragma unroll
            for (int u = 0; u < 4; ++u) {
                sm += v[u];
                sq += v[u] * v[u];
            }
        }
        for (; q < e; ++q) {
            const double v = static_cast<double>(staged ? sval[q] : tval[base + cbase + q]);
            sm += v;
            sq += v * v;
        }
        const int64_t o = static_cast<int64_t>(t) * n_vars + g0 + tid;
        psum[o] = sm;
        psq[o] = sq;
    }
}

// Column statistics from per-tile partials: out[r][v] = sum over the tiles
// of range r (fixed order); a second call with one range finishes the sums.
__global__ void __launch_bounds__(256) k_stats_ranges(const double *__restrict__ psum, const double *__restrict__ psq,
                                                      int n_rows, int rows_per_range, int64_t n_vars,
                                                      double *__restrict__ osum, double *__restrict__ osq) {
    const int64_t v = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (v >= n_vars) return;
    const int r0 = blockIdx.y * rows_per_range;
    const int r1 = r0 + rows_per_range < n_rows ? r0 + rows_per_range : n_rows;
    double s = 0.0, q = 0.0;
    for (int r = r0; r < r1; ++r) {
        s += psum[static_cast<int64_t>(r) * n_vars + v];
        q += psq[static_cast<int64_t>(r) * n_vars + v];
    }
    osum[static_cast<int64_t>(blockIdx.y) * n_vars + v] = s;
    osq[static_cast<int64_t>(blockIdx.y) * n_vars + v] = q;
}

// Entries per gene from the tile-CSC run bounds: cnt[c] += sum over the
// block's tile range of tptr[t][c+1] - tptr[t][c] (coalesced over genes;
// integer atomics, so the totals do not depend on the order). cnt must start
// at zero.
__global__ void __launch_bounds__(256) k_gene_counts_tcsc(const uint32_t *__restrict__ tptr, int n_tiles,
                                                          int tiles_per_block, int64_t n_vars,
                                                          uint32_t *__restrict__ cnt) {
    const int64_t c = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (c >= n_vars) return;
    const int t0 = blockIdx.y * tiles_per_block;
    const int t1 = t0 + tiles_per_block < n_tiles ? t0 + tiles_per_block : n_tiles;
    uint32_t s = 0u;
    for (int t = t0; t < t1; ++t) {
        const uint32_t *tp = tptr + static_cast<int64_t>(t) * (n_vars + 1);
        s += tp[c + 1] - tp[c];
    }
    if (s) atomicAdd(&cnt[c], s);
}

// entries of padding after trow / tval / tpk (read-ahead of the bf16 kernels)
constexpr int64_t kEntryPad = 256;

// ---------------------------------------------------------------------------
// Transposed SpMM with genes owned per 16-lane group (cpu.rs:1282-1389).
//
// Z rows stay in registers, so a wave can only hold a few genes; if the four
// groups of a wave shared each gene, a 1024-thread block would cover 16 GW
// genes and every Q tile would be staged once per 128 genes, with a wave's
// share of a tile ~160 entries between two block barriers. Instead each
// group owns its own GW genes: lane 16g+t holds entry t of group g's current
// run, and the row broadcast (DPP row_newbcast:t) hands every group its own
// entry, so the same gather steps serve four genes at once. A wave covers
// 4 GW genes, a block 64 GW (512 at KP = 64), each Q tile is staged once per
// 512 genes, and a wave does ~650 entries per tile. The four genes
// of a slot (gperm deals genes to groups by entry count, snake order, so
// they have neighbouring counts) run together for the longest of their four
// runs; the shorter runs' lanes carry x = 0 and offset 0. Each group's Z
// rows stay whole in its own lanes (no cross-group sum), accumulated in run
// order: rows ascending inside a tile, tiles ascending.
//
// Entry loads: a run's first 16 entries per group (P0 x 16 for slot 0, the
// densest genes) are loaded one tile ahead into the gather layout; longer
// runs stream through a 64-entry ring per group, the next chunk issued
// before the current one is gathered. Every load index is clamped into the
// group's run (or the run start), so no load leaves the tile-CSC and the
// loads are unconditional. Per tile the Q tile is staged between two block
// barriers; its loads are spread over the slots of the previous tile.
// ---------------------------------------------------------------------------

// Values computed into registers at this point: loads that follow stay after
// it, so a prefetch register read here can take its next load in place
// (no copy at the loop edge, where a copy of a register with a load in
// flight would wait for every outstanding load).
__device__ __forceinline__ void pin_regs(uint32_t &a, float &b) { asm volatile("" : "+v"(a), "+v"(b)::"memory"); }

// Wave priority by the share of the tile's gather steps already done (done,
// total: wave-uniform): waves that are behind issue first, against the SIMD's
// oldest-first arbitration that otherwise leaves the youngest waves of each
// SIMD last to reach the tile barrier.
__device__ __forceinline__ void progress_priority(int done, int total) {
    if (4 * done < total) __builtin_amdgcn_s_setprio(3);
    else if (2 * done < total) __builtin_amdgcn_s_setprio(2);
    else if (4 * done < 3 * total) __builtin_amdgcn_s_setprio(1);
    else __builtin_amdgcn_s_setprio(0);
}

// lane T of each 16-lane row copied to the whole row
template <int T>
__device__ __forceinline__ uint32_t row_bcast(uint32_t v) {
    uint32_t r;
    asm volatile("s_nop 1\n\tv_mov_b32_dpp %0, %1 row_newbcast:%2 row_mask:0xf bank_mask:0xf bound_ctrl:1"
                 : "=v"(r)
                 : "v"(v), "i"(T));
    return r;
}

// Gather steps S..N-1 of a run held as NC components of 16 entries per group
// (step s: component s >> 4, lane s & 15 of each group). The LDS reads run
// two steps ahead of the FMAs through three rotating buffers, and the step
// count is wave-uniform so the exits are scalar branches; lanes past a run
// carry x = 0 and offset 0 (a valid address), so the look-ahead reads are
// unconditional. A scheduling barrier keeps each look-ahead read ahead of the
// current step's FMAs (otherwise the scheduler sinks it next to its use and
// every step waits for its own read).
template <int S, int N, int E, int NC>
struct GatherRun {
    static __device__ __forceinline__ void run(float (&acc)[E], const uint32_t (&ofs)[NC], const float (&x)[NC],
                                               int steps, const char *lds, uint32_t lane_ofs, float4 (&cur)[E / 4],
                                               float4 (&nx1)[E / 4], float4 (&nx2)[E / 4]) {
        if (S >= steps) return;
        if constexpr (S + 2 < N) gather_load<(S + 2) & 15, E>(nx2, ofs[(S + 2) >> 4], lds, lane_ofs);
        __builtin_amdgcn_sched_barrier(0);
        gather_fma<S & 15, E>(acc, cur, x[S >> 4]);
        GatherRun<S + 1, N, E, NC>::run(acc, ofs, x, steps, lds, lane_ofs, nx1, nx2, cur);
    }
};
template <int N, int E, int NC>
struct GatherRun<N, N, E, NC> {
    static __device__ __forceinline__ void run(float (&)[E], const uint32_t (&)[NC], const float (&)[NC], int,
                                               const char *, uint32_t, float4 (&)[E / 4], float4 (&)[E / 4],
                                               float4 (&)[E / 4]) {}
};

// steps: wave-uniform, 1..16 NC
template <int E, int NC>
__device__ __forceinline__ void gather_run(float (&acc)[E], const uint32_t (&ofs)[NC], const float (&x)[NC],
                                           int steps, const char *lds, uint32_t lane_ofs) {
    float4 b0[E / 4], b1[E / 4], b2[E / 4];
    gather_load<0, E>(b0, ofs[0], lds, lane_ofs);
    gather_load<1, E>(b1, ofs[0], lds, lane_ofs);
    GatherRun<0, 16 * NC, E, NC>::run(acc, ofs, x, steps, lds, lane_ofs, b0, b1, b2);
}

// f(std::integral_constant<int, I>) for I = B .. N-1, unrolled
template <int I, int N, typename F>
__device__ __forceinline__ void static_for(F &&f) {
    if constexpr (I < N) {
        f(std::integral_constant<int, I>{});
        static_for<I + 1, N>(f);
    }
}

// word 3 of a raw buffer descriptor (gfx9 family: 32-bit data format, no swizzle)
constexpr int kBufferRsrcWord3 = 0x00020000;

// n_entries: entries in trow / tval / tpk including their padding. BFE: the
// BF16 sketch form, entries from tpk (x~ rounded to bf16). The kernel works
// on KP columns: columns KP * blockIdx.y .. of panels whose rows hold LD
// floats (Q, the partials and their column sums), so a 128-column panel runs
// as two 64-column halves (gridDim.y = 2).
// A prefetched transposed entry: the row in the tile and x~ (FP32 form), or
// the packed bf16 entry of tpk (BF16 sketch form: bf16(x~) in bits 0-15, the
// row in bits 16-23).
template <bool BFE>
struct TEnt {
    uint8_t r;  // the loaded type: no extension on the loop edge
    float v;
};
template <>
struct TEnt<true> {
    uint32_t e;
};

template <int KP, int LD, int GW, int P0, int P1, int RC, int NW, bool BFE = false>
__global__ void __launch_bounds__(NW * 64) k_spmm_t7(const int64_t *__restrict__ indptr,
                                                 const uint32_t *__restrict__ tptr,
                                                 const uint8_t *__restrict__ trow,
                                                 const float *__restrict__ tval,
                                                 const uint32_t *__restrict__ tpk, int64_t n_entries,
                                                 const float *__restrict__ Q,
                                                 const uint32_t *__restrict__ gperm,
                                                 int nch, int ntr, int tiles_per_tr, int64_t n_obs,
                                                 int64_t n_vars, int n_tiles, float *__restrict__ part,
                                                 double *__restrict__ qsum_part) {
    constexpr int E = KP / 16;
    constexpr int NT = NW * kWave;
    constexpr int QC = KP / kWave;                 // Q columns per lane in the column sums
    constexpr int TQ = kTileRows * KP / 4;         // float4s per Q tile
    constexpr int QPT = TQ / NT;                   // direct loads per lane per Q tile
    constexpr int kAfter = 2 * P0;  // entry loads issued after a Q tile's loads, always
    static_assert(GW >= 2 && GW < 16, "one slot per lane of a group, plus the spare lane GW");
    static_assert(P0 >= 1 && P0 <= 4 && P1 >= 1 && P1 <= 4 && RC >= 1 && RC <= 4, "16 to 64 entries per load");
    static_assert(TQ % NT == 0 && kAfter < 64, "whole Q tiles per block; vmcnt range");
    __shared__ float4 qs[2][TQ];  // two Q tiles: 128 KB (KP=64)
    const int tid = threadIdx.x;
    const int xcd = blockIdx.x & 7, slot = blockIdx.x >> 3;
    const int tr = (slot / nch) * 8 + xcd, gc = slot % nch;
    const int wave = __builtin_amdgcn_readfirstlane(tid / kWave);
    const int lane = tid & 63, g = lane >> 4, l = lane & 15;
    const uint32_t lane_ofs = static_cast<uint32_t>(l * E * 4);
    // lane (g, l < GW) holds the gene of slot l of group g
    const int64_t gs = ((static_cast<int64_t>(gc) * NW + wave) * 4 + g) * GW + l;
    const uint32_t gene = (l < GW) ? gperm[gs] : static_cast<uint32_t>(n_vars);
    const bool gvalid = gene < static_cast<uint64_t>(n_vars);
    const bool spare = l == GW;  // reads the tile's entry count
    const uint32_t ia = gvalid ? gene : 0u;
    const uint32_t ib = gvalid ? gene + 1u : static_cast<uint32_t>(n_vars);
    float z[GW][E];
#pragma unroll
    for (int gi = 0; gi < GW; ++gi)
#pragma unroll
        for (int e = 0; e < E; ++e) z[gi][e] = 0.f;
    const int t_begin = tr * tiles_per_tr;
    const int t_end0 = (t_begin + tiles_per_tr) < n_tiles ? (t_begin + tiles_per_tr) : n_tiles;
    const int t_end = (tr < ntr) ? t_end0 : t_begin;

    auto bounds = [&](int tt, uint32_t &a, uint32_t &b) {
        const uint32_t *tp = tptr + static_cast<int64_t>(tt) * (n_vars + 1);
        const uint32_t a0 = tp[ia], b0 = tp[ib];
        a = gvalid ? a0 : 0u;
        b = (gvalid || spare) ? b0 : 0u;
    };
    // Q tile tt into buffer bi by direct global -> LDS loads: wave w's load j
    // fills float4s (j NW + w) * 64 + lane of the tile, rows past n_obs read
    // row 0 (never gathered: no entry refers to them; the column sums stop at
    // the last row). Written as asm so that the compiler does not count these
    // loads against the LDS reads of the other buffer (it would wait for them
    // before every gather); their completion is awaited explicitly (vmcnt,
    // then the block barrier) before the buffer is read.
    static_assert(LD % KP == 0, "whole column halves");
    const int col0 = KP * static_cast<int>(blockIdx.y);
    constexpr int ld4 = LD / 4;
    const float4 *Qh = reinterpret_cast<const float4 *>(Q + col0);
    auto load_q = [&](int tt, int bi) {
        const int64_t r0 = static_cast<int64_t>(tt) * kTileRows;
        static_assert(KP == 64, "a 64-lane load covers four 16-float4 panel rows");
#pragma unroll
        for (int j = 0; j < QPT; ++j) {
            const int i = (j * NW + wave) * kWave;
            const int64_t row = r0 + 4 * (j * NW + wave) + g;  // tile row (i + lane) / 16
            const float4 *gp = Qh + (row < n_obs ? row : 0) * ld4 + l;
            const uint32_t dst = static_cast<uint32_t>(reinterpret_cast<uintptr_t>(&qs[bi][i]));
            asm volatile("global_load_lds_dwordx4 %0, off" ::"v"(gp), "{m0}"(dst) : "memory");
        }
    };
    // Entry loads through buffer descriptors of a tile's trow / tval (base:
    // the tile's first entry, scalar; records: up to the end of the padded
    // arrays): entry j = l + 16 c of a group's run at lo is the lane offset
    // lo + l plus the immediate 16 c (x4 for tval), so a load costs no
    // address arithmetic. Past a run the lanes read the next runs' entries
    // (masked at use); past the arrays the descriptor's range check returns 0.
    struct TileRs {
        __amdgpu_buffer_rsrc_t row, val;
    };
    auto tile_rs = [&](int64_t tb_) {
        const int64_t left = n_entries - tb_;
        const int64_t rb = left < 0x7fffffff ? left : 0x7fffffff;
        const int64_t vb = 4 * left < 0x7fffffff ? 4 * left : 0x7fffffff;
        if constexpr (BFE) {  // one descriptor: tpk
            const __amdgpu_buffer_rsrc_t pk = __builtin_amdgcn_make_buffer_rsrc(
                const_cast<uint32_t *>(tpk + tb_), 0, static_cast<int>(vb), kBufferRsrcWord3);
            return TileRs{pk, pk};
        } else {
            return TileRs{__builtin_amdgcn_make_buffer_rsrc(const_cast<uint8_t *>(trow + tb_), 0,
                                                            static_cast<int>(rb), kBufferRsrcWord3),
                          __builtin_amdgcn_make_buffer_rsrc(const_cast<float *>(tval + tb_), 0, static_cast<int>(vb),
                                                            kBufferRsrcWord3)};
        }
    };
    auto issue = [&](const TileRs &rs, uint32_t lo_, int c, TEnt<BFE> &pe) {
        const uint32_t o = lo_ + static_cast<uint32_t>(l);
        if constexpr (BFE) {
            pe.e = __builtin_amdgcn_raw_buffer_load_b32(rs.row, static_cast<int>(4 * o + 64 * c), 0, 0);
        } else {
            pe.r = __builtin_amdgcn_raw_buffer_load_b8(rs.row, static_cast<int>(o + 16 * c), 0, 0);
            pe.v = __builtin_bit_cast(float,
                                      __builtin_amdgcn_raw_buffer_load_b32(rs.val, static_cast<int>(4 * o + 64 * c), 0, 0));
        }
    };
    auto prep = [&](const TEnt<BFE> &pe, uint32_t len_, uint32_t j0, uint32_t &ofs, float &x) {
        const bool ok = j0 + static_cast<uint32_t>(l) < len_;
        if constexpr (BFE) {
            ofs = ok ? ((pe.e >> 16) & 255u) * (KP * 4) : 0u;
            x = ok ? __builtin_bit_cast(float, pe.e << 16) : 0.f;
        } else {
            ofs = ok ? static_cast<uint32_t>(pe.r) * (KP * 4) : 0u;
            x = ok ? pe.v : 0.f;
        }
    };
    // prefetched components: P0 for slot 0, P1 for slot 1, one for the others
    constexpr int PM = P0 > P1 ? P0 : P1;
    auto ncomp = [](int gi) { return gi == 0 ? P0 : (gi == 1 ? P1 : 1); };
    TEnt<BFE> pent[GW][PM];
    auto issue_slot = [&](int gi, const TileRs &rs, uint32_t lo_) {
#pragma unroll
        for (int c = 0; c < PM; ++c)
            if (c < ncomp(gi)) issue(rs, lo_, c, pent[gi][c]);
    };
    auto slot_bounds = [&](auto gic, uint32_t wl_, uint32_t wh_, uint32_t &lo_, uint32_t &len_) {
        constexpr int GI = decltype(gic)::value;
        lo_ = row_bcast<GI>(wl_);
        len_ = row_bcast<GI>(wh_) - lo_;
    };

    // 1^T Q (centering term): block gc sums, in f64, the staged Q tile rows r
    // with r mod nch == gc (wave w takes the slice's rows w, w + 16, ...)
    const bool do_qsum = qsum_part != nullptr;
    double qsum[QC];
#pragma unroll
    for (int e = 0; e < QC; ++e) qsum[e] = 0.0;
    // run bounds of this tile (cl, ch), the next (nl, nh) and the one after (wl, wh)
    uint32_t cl = 0u, ch = 0u, nl = 0u, nh = 0u, wl = 0u, wh = 0u;
    int64_t tb = 0;
    if (t_begin < t_end) {
        bounds(t_begin, cl, ch);
        tb = indptr[static_cast<int64_t>(t_begin) * kTileRows];
        load_q(t_begin, 0);
        asm volatile("s_waitcnt vmcnt(0)" ::: "memory");  // the first Q tile, once per block
        const TileRs rs0 = tile_rs(tb);
        static_for<0, GW>([&](auto ic) {
            uint32_t lo_, len_;
            slot_bounds(ic, cl, ch, lo_, len_);
            issue_slot(decltype(ic)::value, rs0, lo_);
        });
        const int tn = t_begin + 1 < t_end ? t_begin + 1 : t_begin;
        bounds(tn, nl, nh);
    }
    for (int t = t_begin; t < t_end; ++t) {
        const int bi = (t - t_begin) & 1;
        const int tn = t + 1 < t_end ? t + 1 : t;  // next tile (the last one re-reads itself)
        const int tnn = tn + 1 < t_end ? tn + 1 : tn;
        const char *qb = reinterpret_cast<const char *>(qs[bi]);
        // gathers address both buffers from the array's base: the buffer
        // offset rides in the lane's column offset (one add per tile, none per
        // gather step)
        const char *q0 = reinterpret_cast<const char *>(&qs[0][0]);
        const uint32_t lane_ofs_t = lane_ofs + static_cast<uint32_t>(bi) * static_cast<uint32_t>(sizeof(qs[0]));
        // this tile's Q loads were followed by at least kAfter entry loads:
        // with at most kAfter loads outstanding they have landed in LDS
        asm volatile("s_waitcnt vmcnt(%0)" ::"i"(kAfter) : "memory");
        __syncthreads();  // every wave's share of tile t is in qs[bi]; tile t - 1 (qs[bi ^ 1]) is done
        bounds(tnn, wl, wh);
        if (do_qsum) {  // rows gc + nch w + nch NW i, four reads in flight, summed in row order
            const float *qf = reinterpret_cast<const float *>(qb);
            const int64_t r0 = static_cast<int64_t>(t) * kTileRows;
            const int nr = (n_obs - r0) < kTileRows ? static_cast<int>(n_obs - r0) : kTileRows;
            const int step = nch * NW;
            for (int rr = gc + nch * wave; rr < nr; rr += 4 * step) {
                float v[4][QC];
#pragma unroll
                for (int u = 0; u < 4; ++u)
#pragma unroll
                    for (int e = 0; e < QC; ++e) {
                        const int r = rr + u * step;
                        v[u][e] = qf[(r < nr ? r : 0) * KP + lane + kWave * e];
                    }
#pragma unroll
                for (int u = 0; u < 4; ++u)
#pragma unroll
                    for (int e = 0; e < QC; ++e)
                        if (rr + u * step < nr) qsum[e] += static_cast<double>(v[u][e]);
            }
        }
        const uint32_t tot = static_cast<uint32_t>(__builtin_amdgcn_readlane(static_cast<int>(ch), GW));
        const int64_t ntb = tn != t ? tb + tot : tb;
        const TileRs rs = tile_rs(tb), nrs = tile_rs(ntb);
        // longest run of each slot over the four groups (scalars, by readlane)
        const int rl_len = static_cast<int>(ch - cl);
        int smx[GW];
        int total = 0, done = 0;  // gather steps of the tile
#pragma unroll
        for (int gi = 0; gi < GW; ++gi) {
            const int a0 = __builtin_amdgcn_readlane(rl_len, gi), a1 = __builtin_amdgcn_readlane(rl_len, 16 + gi);
            const int a2 = __builtin_amdgcn_readlane(rl_len, 32 + gi), a3 = __builtin_amdgcn_readlane(rl_len, 48 + gi);
            const int m01 = a0 > a1 ? a0 : a1, m23 = a2 > a3 ? a2 : a3;
            smx[gi] = m01 > m23 ? m01 : m23;
            total += smx[gi];
        }
        // slots from the sparsest to the densest: each slot's prefetch was
        // issued a whole tile earlier, and the next Q tile goes out at the
        // densest slot, whose long gathers cover its latency
        static_for<0, GW>([&](auto ic) {
            constexpr int gi = GW - 1 - decltype(ic)::value;
            constexpr int NC = gi == 0 ? P0 : (gi == 1 ? P1 : 1);
            const std::integral_constant<int, gi> gic{};
            uint32_t lo_, len_, nlo_, nlen_;
            slot_bounds(gic, cl, ch, lo_, len_);
            slot_bounds(gic, nl, nh, nlo_, nlen_);
            const int sm = smx[gi];
            progress_priority(done, total);
            uint32_t ofs[NC];
            float x[NC];
#pragma unroll
            for (int c = 0; c < NC; ++c) {
                prep(pent[gi][c], len_, 16u * c, ofs[c], x[c]);
                pin_regs(ofs[c], x[c]);  // read before the same registers are reloaded
            }
            if (gi == 0 && tn != t) load_q(tn, bi ^ 1);
            issue_slot(gi, nrs, nlo_);
            // the ring: entries from 16 NC on, 16 RC per group per chunk
            TEnt<BFE> rent[RC];
            if (sm > 16 * NC) {
#pragma unroll
                for (int c = 0; c < RC; ++c) issue(rs, lo_, NC + c, rent[c]);
            }
            if (sm > 0) gather_run<E, NC>(z[gi], ofs, x, sm < 16 * NC ? sm : 16 * NC, q0, lane_ofs_t);
            done += sm < 16 * NC ? sm : 16 * NC;
            for (int o = 16 * NC; o < sm; o += 16 * RC) {
                progress_priority(done, total);
                done += sm - o < 16 * RC ? sm - o : 16 * RC;
                uint32_t ofr[RC];
                float xr[RC];
#pragma unroll
                for (int c = 0; c < RC; ++c) {
                    prep(rent[c], len_, static_cast<uint32_t>(o + 16 * c), ofr[c], xr[c]);
                    pin_regs(ofr[c], xr[c]);
                }
                if (o + 16 * RC < sm) {
#pragma unroll
                    for (int c = 0; c < RC; ++c) issue(rs, lo_ + static_cast<uint32_t>(o + 16 * RC), c, rent[c]);
                }
                const int left = sm - o;
                gather_run<E, RC>(z[gi], ofr, xr, left < 16 * RC ? left : 16 * RC, q0, lane_ofs_t);
            }
        });
        cl = nl;
        ch = nh;
        nl = wl;
        nh = wh;
        tb = ntb;
    }
    // every group writes its own genes' rows (16 lanes x E columns)
    static_for<0, GW>([&](auto ic) {
        constexpr int I = decltype(ic)::value;
        const uint32_t c = row_bcast<I>(gene);
        if (tr < ntr && c < static_cast<uint64_t>(n_vars)) {
            float4 *dst = reinterpret_cast<float4 *>(part + (static_cast<int64_t>(tr) * n_vars + c) * LD + col0 + l * E);
#pragma unroll
            for (int e = 0; e < E / 4; ++e)
                dst[e] = make_float4(z[I][4 * e], z[I][4 * e + 1], z[I][4 * e + 2], z[I][4 * e + 3]);
        }
    });
    if (do_qsum) {  // the waves' column sums, added in wave order
        __syncthreads();
        double *red = reinterpret_cast<double *>(qs[0]);  // NW x KP f64 fits in a Q tile buffer
#pragma unroll
        for (int e = 0; e < QC; ++e) red[wave * KP + lane + kWave * e] = qsum[e];
        __syncthreads();
        if (tid < KP) {
            double t = 0.0;
            for (int w = 0; w < NW; ++w) t += red[w * KP + tid];
            qsum_part[(static_cast<int64_t>(tr) * nch + gc) * LD + col0 + tid] = t;
        }
    }
}

// ---------------------------------------------------------------------------
// Forward SpMM with rows owned per 16-lane group (cpu.rs:1185-1229,1471-1523):
// Y[r][:] = sum x~ M[c][:] - mc (+ b_r cs for a PFlog matrix).
//
// Y rows stay in registers; if the four groups of a wave shared each row, a
// 16-wave block would cover 256 rows and stage all of M (four 128 KB gene
// chunks) once per 256 rows: 2 KB of M per row, more than the row's own
// entries. Instead each group owns R rows (lane 16g+t holds entry t of the
// group's current window; the row broadcast hands every group its own
// entry), so a wave covers 4R = 64 rows and a block 1024: M is staged once
// per 1024 rows. A wave sorts its 64 rows by length at the start
// and deals them so that the four rows of a slot have neighbouring lengths
// (the slot runs for the longest of its four windows). Per 512-gene chunk the
// M rows are copied into LDS by direct global -> LDS loads between two block
// barriers; each row's window of the chunk (split offsets) is gathered in
// 16-entry components: the first PC components of a slot's four windows are
// loaded PF slots ahead (across chunk boundaries), longer windows stream the
// rest through a ring. Entry loads go through buffer descriptors based at the
// wave's first entry (32-bit lane offsets, range-checked).
// ---------------------------------------------------------------------------
// The kernel works on KP columns: columns KP * blockIdx.y .. of panels whose
// rows hold LD floats (M, Y, mc, cs), so a 128-column panel runs as two
// 64-column halves (gridDim.y = 2).
// A prefetched forward entry: the gene mod kChunkMod and x~ (FP32 form), or
// the packed bf16 entry of xpk (BF16 sketch form: gene mod kChunkMod in bits
// 0-15, bf16(x~) in bits 16-31).
template <bool BFE>
struct FEnt {
    uint16_t i;  // the loaded type: no extension on the loop edge
    float v;
};
template <>
struct FEnt<true> {
    uint32_t e;
};

// BFE: the BF16 sketch form, entries from xpk.
template <int KP, int LD, int R, int NW, int CG, int PF, int PC, int RC, bool OFFSET = false, bool BFE = false>
__global__ void __launch_bounds__(NW * 64) k_spmm_fwd4(const int64_t *__restrict__ indptr,
                                                       const uint16_t *__restrict__ lg,
                                                       const float *__restrict__ xt,
                                                       const uint32_t *__restrict__ xpk,
                                                       const float *__restrict__ M,
                                                       const float *__restrict__ mc,
                                                       float *__restrict__ Y, const uint32_t *__restrict__ spl,
                                                       int nsplit, int cpc, int nch, int64_t n_obs,
                                                       int64_t n_vars, int64_t nnz,
                                                       const double *__restrict__ rowb,
                                                       const float *__restrict__ cs) {
    constexpr int E = KP / 16;
    constexpr int NT = NW * kWave;
    constexpr int TM = CG * KP / 4;  // float4s per M chunk
    constexpr int MPT = TM / NT;     // direct loads per lane per chunk
    static_assert(R == 16, "one row per lane: a wave's 64 rows are sorted across its lanes");
    static_assert(R % PF == 0 && PF >= 1, "ring slots are static: slot = r % PF in every chunk");
    static_assert(PC >= 1 && PC <= 4 && RC >= 1 && RC <= 4, "16 to 64 entries per group and load");
    static_assert((CG & (CG - 1)) == 0 && CG <= static_cast<int>(kChunkMod), "lg holds gene mod kChunkMod");
    static_assert(TM % NT == 0, "whole float4s per lane");
    __shared__ float4 ms[TM];
    const int tid = threadIdx.x;
    const int wave = __builtin_amdgcn_readfirstlane(tid / kWave);
    const int lane = tid & 63, g = lane >> 4, l = lane & 15;
    const uint32_t lane_ofs = static_cast<uint32_t>(l * E * 4);
    const int64_t row0 = (static_cast<int64_t>(blockIdx.x) * NW + wave) * (4 * R);
    float acc[R][E];
#pragma unroll
    for (int r = 0; r < R; ++r)
#pragma unroll
        for (int e = 0; e < E; ++e) acc[r][e] = 0.f;

    // --- the wave's rows by length, descending (ties: row order); rank k
    // goes to slot k / 4 of group k % 4, i.e. lane (g, s) takes rank 4 s + g
    uint32_t key;
    {
        const int64_t rw = row0 + lane;
        const int64_t rc = rw < n_obs ? rw : n_obs - 1;
        const uint32_t len = rw < n_obs ? static_cast<uint32_t>(indptr[rc + 1] - indptr[rc]) : 0u;
        key = (len << 6) | static_cast<uint32_t>(63 - lane);  // len < 2^26
#pragma unroll
        for (int k = 2; k <= 64; k <<= 1)
#pragma unroll
            for (int j = k >> 1; j > 0; j >>= 1) {
                const uint32_t other = static_cast<uint32_t>(__shfl_xor(static_cast<int>(key), j));
                const bool desc = (lane & k) == 0;  // descending runs, so the whole sequence ends descending
                const bool lower = (lane & j) == 0;
                const uint32_t hi = key > other ? key : other, lo = key > other ? other : key;
                key = (desc == lower) ? hi : lo;
            }
    }
    // the lane's row: row0 + rid (rid < 64); its first entry relative to the wave's
    const uint32_t rid = 63u - (static_cast<uint32_t>(__shfl(static_cast<int>(key), 4 * l + g)) & 63u);
    const int64_t wbase = indptr[row0 < n_obs ? row0 : n_obs];  // the wave's first entry (scalar)
    uint32_t rrel;  // ~0u: no row
    {
        const int64_t row = row0 + rid;
        rrel = row < n_obs ? static_cast<uint32_t>(indptr[row] - wbase) : ~0u;
    }
    auto bounds = [&](int ch, uint32_t &a, uint32_t &b) {  // the row's window of chunk ch, wave-relative
        const int s_lo = ch * cpc, s_hi = (ch + 1) * cpc < nsplit ? (ch + 1) * cpc : nsplit;
        const bool ok = rrel != ~0u;
        const uint32_t *splr = spl + (ok ? row0 + rid : 0) * (nsplit + 1);
        const uint32_t a0 = splr[s_lo], b0 = splr[s_hi];
        a = ok ? rrel + a0 : 0u;
        b = ok ? rrel + b0 : 0u;
    };
    // buffer descriptors of lg / xt from the wave's first entry
    const int64_t left = nnz - wbase;
    const int64_t lb = 2 * left < 0x7fffffff ? 2 * left : 0x7fffffff;
    const int64_t vb = 4 * left < 0x7fffffff ? 4 * left : 0x7fffffff;
    const __amdgpu_buffer_rsrc_t rs_lg =
        BFE ? __builtin_amdgcn_make_buffer_rsrc(const_cast<uint32_t *>(xpk + wbase), 0, static_cast<int>(vb),
                                                kBufferRsrcWord3)
            : __builtin_amdgcn_make_buffer_rsrc(const_cast<uint16_t *>(lg + wbase), 0, static_cast<int>(lb),
                                                kBufferRsrcWord3);
    const __amdgpu_buffer_rsrc_t rs_xt =
        __builtin_amdgcn_make_buffer_rsrc(const_cast<float *>(xt + wbase), 0, static_cast<int>(vb), kBufferRsrcWord3);
    auto issue = [&](uint32_t lo_, int c, FEnt<BFE> &pe) {
        const uint32_t o = lo_ + static_cast<uint32_t>(l);
        if constexpr (BFE) {
            pe.e = __builtin_amdgcn_raw_buffer_load_b32(rs_lg, static_cast<int>(4 * o + 64 * c), 0, 0);
        } else {
            pe.i = __builtin_amdgcn_raw_buffer_load_b16(rs_lg, static_cast<int>(2 * o + 32 * c), 0, 0);
            pe.v = __builtin_bit_cast(float,
                                      __builtin_amdgcn_raw_buffer_load_b32(rs_xt, static_cast<int>(4 * o + 64 * c), 0, 0));
        }
    };
    auto prep = [&](const FEnt<BFE> &pe, uint32_t len_, uint32_t j0, uint32_t &ofs, float &x) {
        const bool ok = j0 + static_cast<uint32_t>(l) < len_;
        if constexpr (BFE) {
            ofs = ok ? (pe.e & static_cast<uint32_t>(CG - 1)) * (KP * 4) : 0u;
            x = ok ? __builtin_bit_cast(float, pe.e & 0xffff0000u) : 0.f;
        } else {
            ofs = ok ? (static_cast<uint32_t>(pe.i) & static_cast<uint32_t>(CG - 1)) * (KP * 4) : 0u;
            x = ok ? pe.v : 0.f;
        }
    };
    auto win = [&](auto rc_, uint32_t a, uint32_t b, uint32_t &lo_, uint32_t &len_) {
        constexpr int RR = decltype(rc_)::value;
        lo_ = row_bcast<RR>(a);
        len_ = row_bcast<RR>(b) - lo_;
    };
    // the ring: PC components of the four windows of a slot, PF slots ahead
    FEnt<BFE> pent[PF][PC];
    auto issue_win = [&](int slot, uint32_t lo_) {
#pragma unroll
        for (int c = 0; c < PC; ++c) issue(lo_, c, pent[slot][c]);
    };
    static_assert(LD % KP == 0, "whole column halves");
    const int col0 = KP * static_cast<int>(blockIdx.y);
    constexpr int ld4 = LD / 4;
    const float4 *Mh = reinterpret_cast<const float4 *>(M + col0);
    static_assert(KP == 64, "a 64-lane load covers four 16-float4 panel rows");
    auto load_m = [&](int ch) {  // wave w's load j fills float4s (j NW + w) * 64 + lane of the chunk
        const int64_t g0 = static_cast<int64_t>(ch) * CG;
#pragma unroll
        for (int j = 0; j < MPT; ++j) {
            const int i = (j * NW + wave) * kWave;
            const int64_t gene = g0 + 4 * (j * NW + wave) + g;  // chunk row (i + lane) / 16
            const float4 *gp = Mh + (gene < n_vars ? gene : 0) * ld4 + l;  // genes past n_vars: never gathered
            const uint32_t dst = static_cast<uint32_t>(reinterpret_cast<uintptr_t>(&ms[i]));
            asm volatile("global_load_lds_dwordx4 %0, off" ::"v"(gp), "{m0}"(dst) : "memory");
        }
    };

    uint32_t ca, cb, na = 0u, nb = 0u;  // windows of this chunk, of the next
    bounds(0, ca, cb);
    static_for<0, PF>([&](auto ic) {
        uint32_t lo_, len_;
        win(ic, ca, cb, lo_, len_);
        issue_win(decltype(ic)::value, lo_);
    });
    for (int ch = 0; ch < nch; ++ch) {
        bounds(ch + 1 < nch ? ch + 1 : ch, na, nb);
        __syncthreads();  // every wave is done with the previous chunk's M rows
        load_m(ch);
        asm volatile("s_waitcnt vmcnt(0)" ::: "memory");
        __syncthreads();  // the chunk's M rows are in LDS
        const char *mb = reinterpret_cast<const char *>(ms);
        // longest window of each slot over the four groups (lane l: slot l)
        uint32_t mx = cb - ca;
        {
            const uint32_t m16 = static_cast<uint32_t>(__shfl_xor(static_cast<int>(mx), 16));
            mx = m16 > mx ? m16 : mx;
            const uint32_t m32 = static_cast<uint32_t>(__shfl_xor(static_cast<int>(mx), 32));
            mx = m32 > mx ? m32 : mx;
        }
        int total = 0, done = 0;
#pragma unroll
        for (int r = 0; r < R; ++r) total += __builtin_amdgcn_readlane(static_cast<int>(mx), r);
        static_for<0, R>([&](auto ic) {
            constexpr int r = decltype(ic)::value;
            uint32_t lo_, len_;
            win(ic, ca, cb, lo_, len_);
            const int sm = __builtin_amdgcn_readlane(static_cast<int>(mx), r);
            progress_priority(done, total);
            uint32_t ofs[PC];
            float x[PC];
#pragma unroll
            for (int c = 0; c < PC; ++c) {
                prep(pent[r % PF][c], len_, 16u * c, ofs[c], x[c]);
                pin_regs(ofs[c], x[c]);  // read before the same registers are reloaded
            }
            {  // the ring slot's next window: slot r + PF of this chunk, or of the next
                uint32_t lo2, len2;
                if constexpr (r + PF < R) {
                    win(std::integral_constant<int, r + PF>{}, ca, cb, lo2, len2);
                } else {
                    win(std::integral_constant<int, r + PF - R>{}, na, nb, lo2, len2);
                }
                issue_win(r % PF, lo2);
            }
            FEnt<BFE> rent[RC];
            if (sm > 16 * PC) {
#pragma unroll
                for (int c = 0; c < RC; ++c) issue(lo_ + 16 * PC, c, rent[c]);
            }
            if (sm > 0) gather_run<E, PC>(acc[r], ofs, x, sm < 16 * PC ? sm : 16 * PC, mb, lane_ofs);
            done += sm < 16 * PC ? sm : 16 * PC;
            for (int o = 16 * PC; o < sm; o += 16 * RC) {
                progress_priority(done, total);
                done += sm - o < 16 * RC ? sm - o : 16 * RC;
                uint32_t ofr[RC];
                float xr[RC];
#pragma unroll
                for (int c = 0; c < RC; ++c) {
                    prep(rent[c], len_, static_cast<uint32_t>(o + 16 * c), ofr[c], xr[c]);
                    pin_regs(ofr[c], xr[c]);
                }
                if (o + 16 * RC < sm) {
#pragma unroll
                    for (int c = 0; c < RC; ++c) issue(lo_ + static_cast<uint32_t>(o + 16 * RC), c, rent[c]);
                }
                const int lft = sm - o;
                gather_run<E, RC>(acc[r], ofr, xr, lft < 16 * RC ? lft : 16 * RC, mb, lane_ofs);
            }
        });
        ca = na;
        cb = nb;
    }
    // every group writes its own rows (16 lanes x E columns)
    static_for<0, R>([&](auto ic) {
        constexpr int r = decltype(ic)::value;
        const int64_t row = row0 + row_bcast<r>(rid);
        if (row < n_obs) {
            float4 *yr = reinterpret_cast<float4 *>(Y + row * LD + col0 + l * E);
            const float br = OFFSET ? static_cast<float>(rowb[row]) : 0.f;
            const float *mch = mc ? mc + col0 : nullptr;
            const float *csh = OFFSET ? cs + col0 : nullptr;
#pragma unroll
            for (int e = 0; e < E / 4; ++e) {
                float4 o;
                o.x = acc[r][4 * e + 0] - (mch ? mch[l * E + 4 * e + 0] : 0.f);
                o.y = acc[r][4 * e + 1] - (mch ? mch[l * E + 4 * e + 1] : 0.f);
                o.z = acc[r][4 * e + 2] - (mch ? mch[l * E + 4 * e + 2] : 0.f);
                o.w = acc[r][4 * e + 3] - (mch ? mch[l * E + 4 * e + 3] : 0.f);
                if constexpr (OFFSET) {
                    o.x += br * csh[l * E + 4 * e + 0];
                    o.y += br * csh[l * E + 4 * e + 1];
                    o.z += br * csh[l * E + 4 * e + 2];
                    o.w += br * csh[l * E + 4 * e + 3];
                }
                yr[e] = o;
            }
        }
    });
}

// ---------------------------------------------------------------------------
// Covariance route (cpu.rs:1663-1854): G = X~^T X~ from the tile-CSC.
// Block (pair p = (I, J), I <= J, of 64-gene blocks; tile range tr) has 16
// waves. Per 256-row tile the slices X~[tile, I] and X~[tile, J] are
// densified in LDS (row-major 256 x 64 f32, zero-filled, then each entry
// scattered: the slice of 64 consecutive genes is one contiguous range of the
// gene-major tile-CSC, and an entry finds its gene by a binary search over the
// 65 run starts staged in LDS). Wave w computes the 32 x 32 quadrant
// (w & 3) of G_IJ over rows [64 (w >> 2), 64 (w >> 2) + 64) of the tile
// with v_mfma_f32_32x32x2f32 (lane l reads A[k0 + l/32][m0 + l%32] and
// B[k0 + l/32][n0 + l%32]: a row-major panel with 64-float rows is
// conflict-free), and adds the tile's f32 result into f64 accumulators. The
// four row quarters are added at the end (in quarter order) and the block
// writes one f64 partial per pair and tile range. f32 products and f32 sums
// over one tile's rows, f64 across tiles (the reference accumulates in f64
// throughout; the difference is ~1e-7 relative per entry).
// ---------------------------------------------------------------------------
constexpr int kCovGB = 64;  // genes per covariance block

__global__ void __launch_bounds__(1024) k_cov_gram(const int64_t *__restrict__ indptr,
                                                  const uint32_t *__restrict__ tptr,
                                                  const uint8_t *__restrict__ trow,
                                                  const float *__restrict__ tval, const int2 *__restrict__ pairs,
                                                  int npairs, int tiles_per_tr, int n_tiles, int64_t n_obs,
                                                  int64_t n_vars, double *__restrict__ part) {
    __shared__ float pan[2][kTileRows * kCovGB];  // 2 x 64 KB
    __shared__ uint32_t rs[2][kCovGB + 1];         // run starts of the two slices (tile-relative)
    const int p = blockIdx.x, tr = blockIdx.y;
    const int2 IJ = pairs[p];
    const bool diag = IJ.x == IJ.y;
    const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
    const int64_t I0 = static_cast<int64_t>(IJ.x) * kCovGB, J0 = static_cast<int64_t>(IJ.y) * kCovGB;
    const int t0 = tr * tiles_per_tr, t1 = (t0 + tiles_per_tr) < n_tiles ? (t0 + tiles_per_tr) : n_tiles;
    const int q = wave & 3, quarter = wave >> 2;  // output quadrant, row quarter of the tile
    const int m0 = 32 * (q >> 1), n0 = 32 * (q & 1);
    const float *pa = pan[0];
    const float *pb = diag ? pan[0] : pan[1];
    v16f acc;
    double acc64[16];
#pragma unroll
    for (int i = 0; i < 16; ++i) acc64[i] = 0.0;
    for (int t = t0; t < t1; ++t) {
        const uint32_t *tp = tptr + static_cast<int64_t>(t) * (n_vars + 1);
        const int64_t base = indptr[static_cast<int64_t>(t) * kTileRows];
        // zero the panels, stage the run starts (genes past n_vars: empty runs)
        float4 *pz = reinterpret_cast<float4 *>(&pan[0][0]);
        const int nz = (diag ? 1 : 2) * kTileRows * kCovGB / 4;
        for (int i = tid; i < nz; i += 1024) pz[i] = make_float4(0.f, 0.f, 0.f, 0.f);
        if (tid < 2 * (kCovGB + 1)) {
            const int s = tid / (kCovGB + 1), c = tid % (kCovGB + 1);
            const int64_t g = (s == 0 ? I0 : J0) + c;
            rs[s][c] = tp[g < n_vars ? g : n_vars];
        }
        __syncthreads();
        // scatter the slices' entries (gene of entry e: largest c with rs[c] <= e)
        for (int s = 0; s < (diag ? 1 : 2); ++s) {
            const uint32_t a = rs[s][0], b = rs[s][kCovGB];
            float *pw = pan[s];
            for (uint32_t e = a + static_cast<uint32_t>(tid); e < b; e += 1024u) {
                int lo = 0;
#pragma unroll
                for (int step = 32; step > 0; step >>= 1) lo = (rs[s][lo + step] <= e) ? lo + step : lo;
                pw[static_cast<int>(trow[base + e]) * kCovGB + lo] = tval[base + e];
            }
        }
        __syncthreads();
        // this wave's quadrant over its 128 rows of the tile
#pragma unroll
        for (int i = 0; i < 16; ++i) acc[i] = 0.f;
        const int rbase = 64 * quarter;
#pragma unroll 8
        for (int k0 = 0; k0 < 64; k0 += 2) {
            const int row = rbase + k0 + (lane >> 5);
            const float av = pa[row * kCovGB + m0 + (lane & 31)];
            const float bv = pb[row * kCovGB + n0 + (lane & 31)];
            acc = __builtin_amdgcn_mfma_f32_32x32x2f32(av, bv, acc, 0, 0, 0);
        }
#pragma unroll
        for (int i = 0; i < 16; ++i) acc64[i] += static_cast<double>(acc[i]);
        __syncthreads();  // the panels are rewritten for the next tile
    }
    // add the row quarters in order: waves 4..15 hand their f64 sums to 0..3
    double *red = reinterpret_cast<double *>(&pan[0][0]);  // 3 x 4 x 64 x 16 f64 = 96 KB
    if (quarter > 0) {
#pragma unroll
        for (int i = 0; i < 16; ++i) red[(((quarter - 1) * 4 + q) * 64 + lane) * 16 + i] = acc64[i];
    }
    __syncthreads();
    if (quarter == 0) {
        double *dst = part + (static_cast<int64_t>(tr) * npairs + p) * (kCovGB * kCovGB);
#pragma unroll
        for (int i = 0; i < 16; ++i) {
            // D[m][n] of v_mfma_f32_32x32x2f32: n = lane % 32, m = 8 (i / 4) + 4 (lane / 32) + i % 4
            const int m = m0 + 8 * (i >> 2) + 4 * (lane >> 5) + (i & 3), n = n0 + (lane & 31);
            dst[m * kCovGB + n] = ((acc64[i] + red[((0 * 4 + q) * 64 + lane) * 16 + i]) +
                                   red[((1 * 4 + q) * 64 + lane) * 16 + i]) +
                                  red[((2 * 4 + q) * 64 + lane) * 16 + i];
        }
    }
}

// cov (n_vars x n_vars f64, both triangles) from the Gram partials: the sum
// over tile ranges in range order, then (G - n mu_i mu_j) / (n - 1)
// (finalize_covariance, cpu.rs:303-322; mu = NULL: G / (n - 1)).
__global__ void __launch_bounds__(256) k_cov_finalize(const double *__restrict__ part, int npairs, int ntr,
                                                      const int *__restrict__ pair_of_blocks, int nblk,
                                                      const double *__restrict__ mu, int64_t n_obs,
                                                      int64_t n_vars, double *__restrict__ cov) {
    const int64_t e = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (e >= n_vars * n_vars) return;
    const int64_t i = e / n_vars, j = e % n_vars;
    const int bi = static_cast<int>(i / kCovGB), bj = static_cast<int>(j / kCovGB);
    const bool sw = bi > bj;  // stored as (min, max) block pair, transposed inside
    const int p = pair_of_blocks[(sw ? bj : bi) * nblk + (sw ? bi : bj)];
    const int li = static_cast<int>((sw ? j : i) % kCovGB), lj = static_cast<int>((sw ? i : j) % kCovGB);
    double g = 0.0;
    for (int r = 0; r < ntr; ++r) g += part[(static_cast<int64_t>(r) * npairs + p) * (kCovGB * kCovGB) + li * kCovGB + lj];
    const double n = static_cast<double>(n_obs);
    const double denom = n - 1.0 > 1.0 ? n - 1.0 : 1.0;
    cov[e] = mu ? (g - n * mu[i] * mu[j]) / denom : g / denom;
}

// Top nc eigenpairs of an ascending EVD of nsel selected pairs (eigenvectors
// in the columns of Z, column-major n x nsel, the largest last):
// variance[pc] = max(w[nsel-1-pc], 0), components row pc = Z[:, nsel-1-pc]
// (f32), and the f32 panel M (n x KP, zero past nc) for the embeddings
// (cpu.rs:1776-1800).
template <int KP>
// c0: the first component of this KP-wide panel (components past 128 come in
// further panels; var is written by the c0 = 0 launch).
__global__ void __launch_bounds__(256) k_cov_topk(const double *__restrict__ Z, const double *__restrict__ w,
                                                  int64_t n, int64_t nsel, int nc, float *__restrict__ comps,
                                                  float *__restrict__ M, double *__restrict__ var, int c0) {
    const int64_t e = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (c0 == 0 && e < nc) {
        const double ev = w[nsel - 1 - e];
        var[e] = ev > 0.0 ? ev : 0.0;
    }
    if (e >= n * KP) return;
    const int64_t v = e / KP;
    const int pc = c0 + static_cast<int>(e % KP);
    const float z = pc < nc ? static_cast<float>(Z[(nsel - 1 - pc) * n + v]) : 0.f;
    M[e] = z;
    if (pc < nc && comps) comps[static_cast<int64_t>(pc) * n + v] = z;
}

// Gram with in-block reduction: each wave accumulates its upper tiles over its
// rows (FP32 MFMA), the four waves of the block add into an LDS tile in a fixed
// order, and the block writes one f64 partial. Two row pairs per iteration
// keep more loads in flight.
template <int KP>
__global__ void __launch_bounds__(256) k_gram2(const float *__restrict__ Y, int64_t nrows,
                                               int64_t rows_per_blk, double *__restrict__ part) {
    constexpr int NT = KP / 32;
    constexpr int NTILE = NT * (NT + 1) / 2;
    __shared__ float gs[KP * KP];
    const int wave = threadIdx.x / kWave, lane = threadIdx.x & 63;
    const int kr = lane >> 5, cl = lane & 31;
    const int64_t b0 = static_cast<int64_t>(blockIdx.x) * rows_per_blk;
    const int64_t b1 = b0 + rows_per_blk < nrows ? b0 + rows_per_blk : nrows;
    v16f acc[NTILE];
#pragma unroll
    for (int t = 0; t < NTILE; ++t) acc[t] = v16f{};
    // waves interleave row pairs: wave w takes pairs w, w+4, ...
    for (int64_t r = b0 + 2 * wave; r < b1; r += 16) {
        float v0[NT], v1[NT];
        const int64_t ra = r + kr, rb = r + 8 + kr;
#pragma unroll
        for (int t = 0; t < NT; ++t) {
            v0[t] = ra < b1 ? Y[ra * KP + t * 32 + cl] : 0.f;
            v1[t] = rb < b1 ? Y[rb * KP + t * 32 + cl] : 0.f;
        }
        int tt = 0;
#pragma unroll
        for (int ti = 0; ti < NT; ++ti)
#pragma unroll
            for (int tj = ti; tj < NT; ++tj) {
                acc[tt] = __builtin_amdgcn_mfma_f32_32x32x2f32(v0[ti], v0[tj], acc[tt], 0, 0, 0);
                acc[tt] = __builtin_amdgcn_mfma_f32_32x32x2f32(v1[ti], v1[tj], acc[tt], 0, 0, 0);
                ++tt;
            }
    }
    for (int w = 0; w < 4; ++w) {
        if (wave == w) {
            int tt = 0;
#pragma unroll
            for (int ti = 0; ti < NT; ++ti)
#pragma unroll
                for (int tj = ti; tj < NT; ++tj) {
#pragma unroll
                    for (int q = 0; q < 16; ++q) {
                        const int i = ti * 32 + 8 * (q / 4) + 4 * kr + (q % 4);
                        const int j = tj * 32 + cl;
                        gs[i * KP + j] = (w == 0 ? 0.f : gs[i * KP + j]) + acc[tt][q];
                    }
                    ++tt;
                }
        }
        __syncthreads();
    }
    double *dst = part + static_cast<int64_t>(blockIdx.x) * KP * KP;
    for (int i = threadIdx.x; i < KP * KP; i += blockDim.x) dst[i] = static_cast<double>(gs[i]);
}


// ---------------------------------------------------------------------------
// Tall f32 panels (KP = 64) on the 16x16x4 f32 MFMA. Operand layout: lane l
// supplies A[l%16][l/16] and B[l/16][l%16]; accumulator q of lane l holds
// D[4(l/16) + q][l%16].
// ---------------------------------------------------------------------------
typedef float v4f __attribute__((ext_vector_type(4)));

// Gram partials G_b = Y_b^T Y_b per block of rows. One float4 load per lane
// covers a 4-row group (lane l: row 4g + l/16, columns 4(l%16) .. +3, 1 KB per
// wave, contiguous), so component e of lane l is A[i][k] = Y[4g + k][4i + e]
// (i = l%16, k = l/16); the 10 products of components ea <= eb give
// D[i][j] = sum_rows Y[row][4i + ea] Y[row][4j + eb] = G[4i + ea][4j + eb],
// which with symmetry covers all of G. f32 accumulation per wave; the 4 waves
// are summed in a fixed order and the block's full 64 x 64 partial is stored
// in f64 (k_sum_parts16 adds the blocks).
template <int KP>
__global__ void __launch_bounds__(256) k_gram4(const float *__restrict__ Y, int64_t nrows, int64_t rows_per_blk,
                                               double *__restrict__ part) {
    static_assert(KP == 64, "one float4 per lane per row");
    constexpr int NTL = 10, U = 4;
    __shared__ float red[4][NTL][256];
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int64_t b0 = static_cast<int64_t>(blockIdx.x) * rows_per_blk;
    const int64_t b1 = b0 + rows_per_blk < nrows ? b0 + rows_per_blk : nrows;
    const int64_t ngrp = b1 > b0 ? (b1 - b0 + 3) / 4 : 0;
    v4f acc[NTL];
#pragma unroll
    for (int t = 0; t < NTL; ++t) acc[t] = v4f{0.f, 0.f, 0.f, 0.f};
    for (int64_t g0 = wave; g0 < ngrp; g0 += 4 * U) {
        float4 v[U];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const int64_t row = b0 + 4 * (g0 + 4 * u) + (lane >> 4);
            const bool ok = row < b1;
            const int64_t rr = ok ? row : b0;  // in bounds; zeroed below
            v[u] = *reinterpret_cast<const float4 *>(Y + rr * KP + 4 * (lane & 15));
            if (!ok) v[u] = make_float4(0.f, 0.f, 0.f, 0.f);
        }
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const float c[4] = {v[u].x, v[u].y, v[u].z, v[u].w};
            int t = 0;
#pragma unroll
            for (int ea = 0; ea < 4; ++ea)
#pragma unroll
                for (int eb = ea; eb < 4; ++eb) {
                    acc[t] = __builtin_amdgcn_mfma_f32_16x16x4f32(c[ea], c[eb], acc[t], 0, 0, 0);
                    ++t;
                }
        }
    }
#pragma unroll
    for (int t = 0; t < NTL; ++t)
#pragma unroll
        for (int q = 0; q < 4; ++q) red[wave][t][lane * 4 + q] = acc[t][q];
    __syncthreads();
    double *dst = part + static_cast<int64_t>(blockIdx.x) * KP * KP;
    for (int e = threadIdx.x; e < KP * KP; e += blockDim.x) {
        const int a = e / KP, b = e % KP;
        int ea = a & 3, eb = b & 3, i = a >> 2, j = b >> 2;
        if (ea > eb) {
            int x = ea; ea = eb; eb = x;
            x = i; i = j; j = x;
        }
        const int t = ea * 4 - ea * (ea - 1) / 2 + (eb - ea);  // index of (ea, eb) in the ea <= eb order
        const int idx = (16 * (i >> 2) + j) * 4 + (i & 3);     // D[i][j]: lane 16(i/4) + j, reg i%4
        dst[e] = static_cast<double>(((red[0][t][idx] + red[1][t][idx]) + red[2][t][idx]) + red[3][t][idx]);
    }
}

// G = sum of nparts KP x KP f64 partials, in a fixed order: block b sums
// entries [16b, 16b + 16), 16 lanes per entry each taking every 16th partial
// (eight loads in flight, added as a fixed tree), then the 16 lane sums are
// added in order.
template <int KP>
__global__ void __launch_bounds__(256) k_sum_parts16(const double *__restrict__ part, int nparts,
                                                     double *__restrict__ G) {
    __shared__ double red[16][17];
    const int e = blockIdx.x * 16 + (threadIdx.x & 15), g = threadIdx.x >> 4;
    constexpr int64_t KK = static_cast<int64_t>(KP) * KP;
    double acc = 0.0;
    if (e < KP * KP) {
        int p = g;
        for (; p + 7 * 16 < nparts; p += 8 * 16) {
            double v[8];
#pragma unroll
            for (int u = 0; u < 8; ++u) v[u] = part[static_cast<int64_t>(p + 16 * u) * KK + e];
            acc += ((v[0] + v[1]) + (v[2] + v[3])) + ((v[4] + v[5]) + (v[6] + v[7]));
        }
        for (; p < nparts; p += 16) acc += part[static_cast<int64_t>(p) * KK + e];
    }
    red[g][threadIdx.x & 15] = acc;
    __syncthreads();
    if (g == 0 && e < KP * KP) {
        double s = 0.0;
#pragma unroll
        for (int h = 0; h < 16; ++h) s += red[h][threadIdx.x & 15];
        G[e] = s;
    }
}

// out = Y W (Y: nrows x 64 f32, W: 64 x 64 f32), the first out_cols columns
// stored with leading dimension out_ld. When out_ld == out_cols (and out is
// 16-byte aligned) a tile's 16 output rows are one contiguous block: it is
// staged in LDS and written as float4s (4 wave-wide stores for 64 columns,
// instead of 16 stores of 64-byte pieces). A wave takes 16-row tiles: lane l
// loads 4 float4 of row r0 + l%16 (columns 16m + 4(l/16) .. +3), so component
// e of load m is A[i][k] for inner index 16m + 4k + e; the matching B values
// W[16m + 4(l/16) + e][16jt + l%16] stay in 64 registers per lane. 64 MFMAs
// per tile, 40 when W is upper triangular (UPPER: the blocks m > jt are zero
// and skipped); lane l then holds out[r0 + 4(l/16) + q][16jt + l%16].
template <int KP, bool UPPER>
__global__ void __launch_bounds__(256) k_apply4(const float *__restrict__ Y, const float *__restrict__ W,
                                                float *__restrict__ out, int64_t nrows, int out_cols,
                                                int64_t out_ld) {
    static_assert(KP == 64, "four float4 per lane per row");
    __shared__ __align__(16) float tile_out[4][16 * KP];  // per wave: 16 rows x out_cols, compact
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int i16 = lane & 15, k4 = lane >> 4;
    const bool packed = out_ld == out_cols && (reinterpret_cast<uintptr_t>(out) & 15u) == 0;
    float *tw = tile_out[wave];
    float wr[4][4][4];
#pragma unroll
    for (int m = 0; m < 4; ++m)
#pragma unroll
        for (int e = 0; e < 4; ++e)
#pragma unroll
            for (int jt = 0; jt < 4; ++jt) wr[m][e][jt] = W[(16 * m + 4 * k4 + e) * KP + 16 * jt + i16];
    const int64_t ntiles = (nrows + 15) / 16;
    for (int64_t tile = static_cast<int64_t>(blockIdx.x) * 4 + wave; tile < ntiles;
         tile += static_cast<int64_t>(gridDim.x) * 4) {
        const int64_t r0 = tile * 16;
        const int64_t row = r0 + i16;
        const bool ok = row < nrows;
        const int64_t rr = ok ? row : nrows - 1;  // in bounds; zeroed below
        float4 y[4];
#pragma unroll
        for (int m = 0; m < 4; ++m) {
            y[m] = *reinterpret_cast<const float4 *>(Y + rr * KP + 16 * m + 4 * k4);
            if (!ok) y[m] = make_float4(0.f, 0.f, 0.f, 0.f);
        }
        v4f acc[4];
#pragma unroll
        for (int jt = 0; jt < 4; ++jt) acc[jt] = v4f{0.f, 0.f, 0.f, 0.f};
#pragma unroll
        for (int m = 0; m < 4; ++m) {
            const float c[4] = {y[m].x, y[m].y, y[m].z, y[m].w};
#pragma unroll
            for (int e = 0; e < 4; ++e)
#pragma unroll
                for (int jt = UPPER ? m : 0; jt < 4; ++jt)
                    acc[jt] = __builtin_amdgcn_mfma_f32_16x16x4f32(c[e], wr[m][e][jt], acc[jt], 0, 0, 0);
        }
        if (packed) {
#pragma unroll
            for (int jt = 0; jt < 4; ++jt) {
                const int col = 16 * jt + i16;
                if (col < out_cols)
#pragma unroll
                    for (int q = 0; q < 4; ++q) tw[(4 * k4 + q) * out_cols + col] = acc[jt][q];
            }
            wave_lds_sync();
            const int64_t left = nrows - r0;
            const int nv = static_cast<int>(left < 16 ? left : 16) * out_cols;  // floats of the tile
            float *dst = out + r0 * out_ld;
            const int n4 = nv >> 2;
            for (int i = lane; i < n4; i += kWave)
                reinterpret_cast<float4 *>(dst)[i] = reinterpret_cast<const float4 *>(tw)[i];
            for (int i = 4 * n4 + lane; i < nv; i += kWave) dst[i] = tw[i];
            wave_lds_sync();  // the tile is read before the next one is staged
        } else {
#pragma unroll
            for (int jt = 0; jt < 4; ++jt) {
                const int col = 16 * jt + i16;
                if (col < out_cols)
#pragma unroll
                    for (int q = 0; q < 4; ++q) {
                        const int64_t orow = r0 + 4 * k4 + q;
                        if (orow < nrows) out[orow * out_ld + col] = acc[jt][q];
                    }
            }
        }
    }
}

// One-sided Jacobi SVD of the k x k R (f64), columns held in LDS for KP <= 64
// (A and V) and A in LDS / V in global scratch for KP = 128.
template <int KP>
__global__ void __launch_bounds__(256) k_jacobi_lds(const double *__restrict__ R, int k, double *__restrict__ Vg,
                                                    double *__restrict__ U, double *__restrict__ V,
                                                    double *__restrict__ sigma, int *__restrict__ status) {
    constexpr bool VLDS = KP <= 64;
    __shared__ double A[KP * KP];                 // column-major: A[c*KP + i]
    __shared__ double Vs[VLDS ? KP * KP : 1];
    __shared__ int changed, large;
    __shared__ double nrm[KP];
    __shared__ int order[KP];
    double *Vc = VLDS ? Vs : Vg;
    const int tid = threadIdx.x;
    const int n = (k + 1) & ~1;
    for (int t = tid; t < KP * KP; t += blockDim.x) {
        const int i = t % KP, c = t / KP;
        A[t] = (i < k && c < k) ? R[i * KP + c] : 0.0;
        Vc[t] = (i == c && c < n) ? 1.0 : 0.0;
    }
    __syncthreads();
    constexpr int TPP = 8;
    const int npairs = n / 2;
    for (int sweep = 0; sweep < 60 && n >= 2; ++sweep) {
        if (tid == 0) changed = large = 0;
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
                if (p > q) { const int sw = p; p = q; q = sw; }
                double *ap = A + p * KP, *aq = A + q * KP;
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
                    const double c = 1.0 / sqrt(1.0 + tt * tt), sn = c * tt;
                    for (int i = sub; i < n; i += TPP) {
                        const double x = ap[i], y = aq[i];
                        ap[i] = c * x - sn * y;
                        aq[i] = sn * x + c * y;
                    }
                    double *vp = Vc + p * KP, *vq = Vc + q * KP;
                    for (int i = sub; i < n; i += TPP) {
                        const double x = vp[i], y = vq[i];
                        vp[i] = c * x - sn * y;
                        vq[i] = sn * x + c * y;
                    }
                    if (sub == 0) changed = 1;
                    if (sub == 0 && fabs(ga) > 1e-5 * sqrt(al * be)) large = 1;
                }
            }
            __syncthreads();
        }
        if (!changed || !large) break;  // a sweep whose pairs were all within 1e-5 leaves ~1e-10
        if (sweep == 59 && tid == 0) atomicOr(status, 2);
        __syncthreads();
    }
    for (int c = tid; c < KP; c += blockDim.x) {
        double s = 0.0;
        if (c < k)
            for (int i = 0; i < n; ++i) s += A[c * KP + i] * A[c * KP + i];
        nrm[c] = c < k ? sqrt(s) : -1.0;
    }
    __syncthreads();
    for (int c = tid; c < KP; c += blockDim.x) {
        int rank = 0;
        for (int d = 0; d < KP; ++d)
            if (nrm[d] > nrm[c] || (nrm[d] == nrm[c] && d < c)) ++rank;
        order[rank] = c;
    }
    __syncthreads();
    for (int t = tid; t < KP * KP; t += blockDim.x) {
        const int i = t / KP, jj = t % KP;
        const int c = order[jj];
        const double sg = nrm[c] > 0.0 ? nrm[c] : 0.0;
        U[t] = (jj < k && sg > 0.0) ? A[c * KP + i] / sg : 0.0;
        V[t] = jj < k ? Vc[c * KP + i] : 0.0;
    }
    for (int jj = tid; jj < KP; jj += blockDim.x) sigma[jj] = jj < k ? (nrm[order[jj]] > 0 ? nrm[order[jj]] : 0.0) : 0.0;
}

// Q = Y W on MFMA with the A operand taken from registers: lane (kr, cl) loads
// row cl of its 32-row tile once and uses columns 2s+kr; W in LDS.
template <int KP>
__global__ void __launch_bounds__(256) k_apply2(const float *__restrict__ Y, const float *__restrict__ W,
                                                float *__restrict__ out, int64_t nrows, int out_cols,
                                                int64_t out_ld) {
    constexpr int NT = KP / 32;
    __shared__ float ws[KP * KP];
    for (int t = threadIdx.x; t < KP * KP; t += blockDim.x) ws[t] = W[t];
    __syncthreads();
    const int wave = threadIdx.x / kWave, lane = threadIdx.x & 63;
    const int kr = lane >> 5, cl = lane & 31;
    const int64_t ntiles = (nrows + 31) / 32;
    for (int64_t tile = static_cast<int64_t>(blockIdx.x) * 4 + wave; tile < ntiles;
         tile += static_cast<int64_t>(gridDim.x) * 4) {
        const int64_t r0 = tile * 32;
        const int64_t row = r0 + cl;
        float yr[KP / 2];
        if (row < nrows) {
            const float2 *src = reinterpret_cast<const float2 *>(Y + row * KP);
#pragma unroll
            for (int s2 = 0; s2 < KP / 2; ++s2) {
                const float2 v = src[s2];
                yr[s2] = kr ? v.y : v.x;
            }
        } else {
#pragma unroll
            for (int s2 = 0; s2 < KP / 2; ++s2) yr[s2] = 0.f;
        }
#pragma unroll
        for (int nt = 0; nt < NT; ++nt) {
            v16f acc = v16f{};
#pragma unroll
            for (int s2 = 0; s2 < KP / 2; ++s2)
                acc = __builtin_amdgcn_mfma_f32_32x32x2f32(yr[s2], ws[(2 * s2 + kr) * KP + nt * 32 + cl], acc, 0, 0, 0);
            const int col = nt * 32 + cl;
            if (col < out_cols) {
#pragma unroll
                for (int q = 0; q < 16; ++q) {
                    const int64_t orow = r0 + 8 * (q / 4) + 4 * kr + (q % 4);
                    if (orow < nrows) out[orow * out_ld + col] = acc[q];
                }
            }
        }
    }
}

// ---------------------------------------------------------------------------
// Cholesky G (+ shift * trace(G) * I) = R^T R and W = R^-1 (f64, one block of
// 256 threads, G in LDS as a padded full matrix). Right-looking with one
// barrier per step: at step j every thread reads the pivot d = a[j][j] (final
// since the previous barrier) and updates its (i, c) entries of the trailing
// upper triangle, a[i][c] -= (a[j][i] / r)(a[j][c] / r) with r = sqrt(d), on a
// 16 x 16 thread grid (no index division); row j itself is scaled in place
// during step j + 1, when no thread reads it any more. W is then solved column
// by column, W[i][c] = -(sum_{l=i+1..c} R[i][l] W[l][c]) / R[i][i], LPC lanes
// per column (partial sums combined by shuffles in a fixed order, within one
// wave, so no barriers), stored transposed in the free lower triangle.
// Pivot rule (both Cholesky kernels): floor_tau == 0 is strict (a pivot that
// is not positive and finite sets status bit 1, as CholeskyQR2's non-SPD
// error); floor_tau > 0 floors a pivot below floor_tau * the mean (shifted)
// diagonal at that value and sets bit 4. R stays nonsingular, so Q = A R^-1
// still spans A: an exactly rank-deficient A gets a floored direction that
// the next pass normalises. The floor is relative to the matrix, not to the
// column: a column of pure round-off is then scaled up by at most
// 1/sqrt(floor_tau) per pass instead of without bound. A non-finite pivot
// still sets bit 1.
// ---------------------------------------------------------------------------
template <int KP>
__global__ void __launch_bounds__(256) k_chol2(const double *__restrict__ G, int k, double shift,
                                               double *__restrict__ R64, double *__restrict__ W64,
                                               float *__restrict__ W32, int *__restrict__ status,
                                               double floor_tau) {
    constexpr int LD = KP + 1;           // padded row; column KP holds W's diagonal
    constexpr int LPC = 256 / KP;        // lanes per column in the inverse (4 or 2)
    __shared__ double a[KP * LD];
    __shared__ double rinv[KP];
    __shared__ double trs;
    const int tid = threadIdx.x, tx = tid & 15, ty = tid >> 4;
    for (int t = tid; t < KP * KP; t += 256) {
        const int i = t / KP, c = t % KP;
        a[i * LD + c] = (i < k && c < k) ? G[t] : 0.0;
    }
    __syncthreads();
    if (tid < 64) {
        double s = 0.0;
        for (int i = tid; i < k; i += 64) s += a[i * LD + i];
        s = wave_sum(s);
        if (tid == 0) trs = s;
    }
    __syncthreads();
    if (shift > 0.0) {
        for (int i = tid; i < k; i += 256) a[i * LD + i] += shift * trs;
        __syncthreads();
    }
    const double mdiag = k > 0 ? trs * (1.0 + (shift > 0.0 ? shift * k : 0.0)) / k : 0.0;
    const double fl = floor_tau * mdiag > 0.0 ? floor_tau * mdiag : 1.0;
    bool bad = false, floored = false;
    double r_prev = 1.0, inv_prev = 1.0;
    for (int j = 0; j < k; ++j) {
        const double d = a[j * LD + j];
        double ok_d;
        if (floor_tau > 0.0) {
            const bool fin = isfinite(d);
            bad |= !fin;
            floored |= fin && !(d > fl);
            ok_d = (fin && d > fl) ? d : (fin ? fl : 1.0);
        } else {
            const bool ok = (d > 0.0) && isfinite(d);
            bad |= !ok;
            ok_d = ok ? d : 1.0;
        }
        const double r = sqrt(ok_d);
        const double inv = 1.0 / r;
        if (j > 0) {  // scale row j-1 (no thread reads it in this step)
            for (int c = j - 1 + tid; c < k; c += 256)
                a[(j - 1) * LD + c] = (c == j - 1) ? r_prev : a[(j - 1) * LD + c] * inv_prev;
        }
        for (int i = j + 1 + ty; i < k; i += 16) {
            const double ri = a[j * LD + i] * inv;
            for (int c = i + ((tx - (i - (j + 1))) & 15); c < k; c += 16) {
                // c runs over the columns >= i congruent to j+1+tx mod 16
                a[i * LD + c] -= ri * (a[j * LD + c] * inv);
            }
        }
        r_prev = r;
        inv_prev = inv;
        __syncthreads();
    }
    if (k > 0)
        for (int c = k - 1 + tid; c < k; c += 256)
            a[(k - 1) * LD + c] = (c == k - 1) ? r_prev : a[(k - 1) * LD + c] * inv_prev;
    __syncthreads();
    for (int i = tid; i < k; i += 256) rinv[i] = 1.0 / a[i * LD + i];
    if (floored && tid == 0) atomicOr(status, 4);
    __syncthreads();
    for (int t = tid; t < KP * KP; t += 256) {
        const int i = t / KP, c = t % KP;
        R64[t] = (i < k && c < k && c >= i) ? a[i * LD + c] : 0.0;
    }
    // inverse: lanes LPC*c .. LPC*c + LPC-1 own column c
    {
        const int c = tid / LPC, q = tid % LPC;
        if (c < k) {
            const double wcc = rinv[c];
            if (q == 0) a[c * LD + KP] = wcc;
            for (int i = c - 1; i >= 0; --i) {
                double part = 0.0;
                for (int l = i + 1 + q; l <= c; l += LPC) part += a[i * LD + l] * (l == c ? wcc : a[c * LD + l]);
#pragma unroll
                for (int o = 1; o < LPC; o <<= 1) part += __shfl_xor(part, o);
                if (q == 0) a[c * LD + i] = -part * rinv[i];
            }
        }
    }
    __syncthreads();
    for (int t = tid; t < KP * KP; t += 256) {
        const int i = t / KP, c = t % KP;
        const double v = (i < k && c < k) ? (c > i ? a[c * LD + i] : (c == i ? a[c * LD + KP] : 0.0)) : 0.0;
        W64[t] = v;
        W32[t] = static_cast<float>(v);
    }
    if (bad && tid == 0) atomicOr(status, 1);
}

// Cholesky factor for KP = 64 in ONE wave, no block barriers: lane c holds
// column c of the Gram in 64 registers (entries with i >= k or c >= k set to
// the identity, so every step runs unconditionally and the padded factor is
// [R 0; 0 I]). Step j: the pivot a[j][j] comes from lane j by readlane,
// r = d rsqrt(d), lane c forms R[j][c] = a[j][c] / r and publishes it to LDS
// (double buffered), and every lane updates a[i][c] -= R[j][i] R[j][c] for
// i > j with R[j][i] read back as a broadcast (measured faster than 63
// readlanes per step). Lanes c < i keep updating their lower-triangle
// registers (never read), which keeps the code branch-free. After a block
// barrier R is inverted (W64 / W32 = R^-1) from its LDS copy in the same
// launch, by 2 x 2 blocks: the two diagonal blocks in 32 steps on one wave,
// the off-diagonal block by two 32 x 32 products on all four waves. The pivot rule is k_chol2's; FLOOR selects the floor (per step
// a max with the floor, non-finite values checked on the input and on the
// final diagonal: per-step compares would cost scalar registers).
// ---------------------------------------------------------------------------
template <bool FLOOR>
__global__ void __launch_bounds__(256) k_chol_w(const double *__restrict__ G, int k, double shift,
                                               double *__restrict__ R64, double *__restrict__ W64,
                                               float *__restrict__ W32, int *__restrict__ status,
                                               double floor_tau) {
    constexpr int KP = 64, H = KP / 2;
    __shared__ double rowb[2][KP];
    __shared__ double Rs[KP * KP];
    __shared__ double rinv[KP];
    __shared__ double Wl[KP * KP];     // R^-1, row-major
    __shared__ double Tl[H * (H + 1)];  // B C^-1 (padded rows)
    // the masked Gram staged row-major in LDS by all four waves (four times
    // the loads in flight of one wave); wave 0 then factors it alone
    for (int t = threadIdx.x; t < KP * KP; t += blockDim.x) {
        const int i = t / KP, cc = t % KP;
        Rs[t] = (i < k && cc < k) ? G[t] : (i == cc ? 1.0 : 0.0);
    }
    __syncthreads();
    const int c = threadIdx.x;
    if (c < KP) {
        const double tr = wave_sum_dpp(c < k ? Rs[c * KP + c] : 0.0);
        const double sh = (shift > 0.0 && c < k) ? shift * tr : 0.0;
        const double mdiag = k > 0 ? (tr + (shift > 0.0 ? shift * tr * k : 0.0)) / k : 0.0;
        const double fl = floor_tau * mdiag > 0.0 ? floor_tau * mdiag : 1.0;  // FLOOR: the pivot floor
        double a[KP];
#pragma unroll
        for (int i = 0; i < KP; ++i) a[i] = Rs[i * KP + c] + ((i == c) ? sh : 0.0);
        wave_lds_sync();
        bool bad = false;
        double rdiag = 1.0;  // FLOOR: R[c][c], kept by lane c
        if constexpr (FLOOR) {
            // non-finite input is checked once here and on the final diagonal, so
            // the step needs no compare (per-step masks cost scalar registers)
#pragma unroll
            for (int i = 0; i < KP; ++i) bad |= !isfinite(a[i]);
        }
#pragma unroll
        for (int j = 0; j < KP; ++j) {
            const double d = readlane_f64(a[j], j);
            double dd;  // select the input, not the result: no branch
            if constexpr (FLOOR) {
                dd = fmax(d, fl);
            } else {
                const bool ok = (d > 0.0) && isfinite(d);
                bad |= !ok;
                dd = ok ? d : 1.0;
            }
            const double ir = rsq_nr(dd);
            const double r = dd * ir;
            const double t = (c == j) ? r : a[j] * ir;
            if constexpr (FLOOR) rdiag = (c == j) ? r : rdiag;
            a[j] = t;
            rowb[j & 1][c] = t;
            wave_lds_sync();
#pragma unroll
            for (int i = j + 1; i < KP; ++i) a[i] -= rowb[j & 1][i] * t;
        }
        if constexpr (FLOOR) {
            // R[c][c] = sqrt(max(pivot, floor)): equal to sqrt(floor) where it floored
            const bool mine = c < k;
            bad |= mine && !isfinite(rdiag);
            const bool floored = mine && rdiag * rdiag <= fl * (1.0 + 1e-12);
            if (__ballot(floored) != 0 && c == 0) atomicOr(status, 4);
            bad = __ballot(bad) != 0;
        }
        if (bad && c == 0) atomicOr(status, 1);
        // R out, and into LDS (the Gram's reads above are done) for the inverse
#pragma unroll
        for (int i = 0; i < KP; ++i) {
            const double v = (i <= c && c < k) ? a[i] : 0.0;
            R64[i * KP + c] = v;
            Rs[i * KP + c] = v;
        }
    }
    __syncthreads();
    // W = R^-1 by 2 x 2 blocks of H = 32: R = [A B; 0 C], W = [A^-1, -A^-1 B C^-1;
    // 0, C^-1]. Wave 0 inverts A (lanes 0-31) and C (lanes 32-63) together,
    // right-looking: lane c holds column c of diag(A, C)^-1; X = I; for
    // j = H-1 .. 0, X[j][:] /= R[j][j], X[i][:] -= R[i][j] X[j][:] for i < j
    // (block-local indices), R[i][j] read from LDS (one address per half-wave).
    // Pivots j >= k are taken as 1, so padded columns never reach columns c < k.
    // Then all four waves form T = B C^-1 and the upper-right block -A^-1 T.
    if (c < KP) {
        rinv[c] = c < k ? 1.0 / Rs[c * KP + c] : 1.0;
        wave_lds_sync();
        const int off = c < H ? 0 : H, cl = c - off;
        double x[H];
#pragma unroll
        for (int i = 0; i < H; ++i) x[i] = (i == cl) ? 1.0 : 0.0;
#pragma unroll
        for (int j = H - 1; j >= 0; --j) {
            const double w = x[j] * rinv[off + j];
            x[j] = w;
#pragma unroll
            for (int i = 0; i < j; ++i) x[i] -= Rs[(off + i) * KP + off + j] * w;
            // a fence per step keeps every step's reads from being hoisted to
            // the top (which would need ~500 registers)
            wave_lds_sync();
        }
#pragma unroll
        for (int i = 0; i < H; ++i) Wl[(off + i) * KP + c] = x[i];
    }
    __syncthreads();
    // T[i][cc] = sum_{q <= cc} B[i][q] C^-1[q][cc]  (C^-1 upper triangular)
    for (int e = threadIdx.x; e < H * H; e += blockDim.x) {
        const int i = e / H, cc = e % H;
        double t = 0.0;
        for (int q = 0; q <= cc; ++q) t = fma(Rs[i * KP + H + q], Wl[(H + q) * KP + H + cc], t);
        Tl[i * (H + 1) + cc] = t;
    }
    __syncthreads();
    // W[i][H + cc] = -sum_{q >= i} A^-1[i][q] T[q][cc]  (A^-1 upper triangular)
    for (int e = threadIdx.x; e < H * H; e += blockDim.x) {
        const int i = e / H, cc = e % H;
        double t = 0.0;
        for (int q = i; q < H; ++q) t = fma(Wl[i * KP + q], Tl[q * (H + 1) + cc], t);
        Wl[i * KP + H + cc] = -t;
    }
    __syncthreads();
    for (int t = threadIdx.x; t < KP * KP; t += blockDim.x) {
        const int i = t / KP, cc = t % KP;
        const double v = (i <= cc && cc < k) ? Wl[t] : 0.0;
        W64[t] = v;
        W32[t] = static_cast<float>(v);
    }
}

// One-sided (Hestenes) Jacobi SVD of the k x k R in f64 for KP = 64: the
// columns of A (= R, rotated) and V live in LDS, column-major. Each round of
// the round-robin ordering pairs all columns; pair slot s (0..31) is one
// 16-lane DPP row, lane l of it holding rows 4l..4l+3 of the pair's columns.
// The three inner products are summed over the row by DPP (identical bits in
// all 16 lanes, so they agree on the rotation), which is
// tan = sign(d) 2g / (|d| + sqrt(d^2 + 4g^2)), d = |y|^2 - |x|^2, g = x.y (the
// usual zeta formula multiplied through by |2g|), cos = rsqrt(1 + tan^2).
// A pair is rotated while |x.y| > 1e-10 |x| |y|: converged columns are then
// orthogonal to 1e-10 (singular values exact to ~1e-20 relative, vectors to
// ~1e-10, far below the f32 bases this SVD finishes); a tighter threshold
// only adds sweeps that rotate by rounding noise. The sweeps stop after one
// whose largest |x.y| / (|x| |y|) was at most 1e-5: Jacobi converges
// quadratically, so its rotations leave ~1e-10 and the sweep that would only
// confirm it is skipped (7 sweeps instead of 9 at 1M x 2k, vectors still
// orthogonal to ~1e-9).
// One barrier per round; the per-sweep "rotated" flags rotate through three
// slots so that a reset never races a read. Outputs as k_jacobi_lds: U, V
// row-major with columns sorted by descending singular value, sigma.
template <int KP>
__global__ void __launch_bounds__(512) k_jacobi_w(const double *__restrict__ R, int k, double *__restrict__ U,
                                                  double *__restrict__ V, double *__restrict__ sigma,
                                                  int *__restrict__ status) {
    static_assert(KP == 64, "16 lanes x 4 rows per column");
    __shared__ __align__(16) double A[KP * KP];
    __shared__ __align__(16) double Vs[KP * KP];
    __shared__ int changed[3], large[3];
    __shared__ double nrm[KP];
    __shared__ int order[KP];
    const int tid = threadIdx.x, l16 = tid & 15, slot = tid >> 4;
    const int n = (k + 1) & ~1;
    for (int t = tid; t < KP * KP; t += blockDim.x) {
        const int i = t % KP, c = t / KP;
        A[t] = (i < k && c < k) ? R[i * KP + c] : 0.0;
        Vs[t] = (i == c && c < n) ? 1.0 : 0.0;
    }
    if (tid < 3) changed[tid] = large[tid] = 0;
    __syncthreads();
    const int npairs = n / 2;
    const bool act = slot < npairs;
    int sweep = 0;
    for (; sweep < 60 && n >= 2; ++sweep) {
        if (tid == 0) changed[(sweep + 1) % 3] = large[(sweep + 1) % 3] = 0;
        for (int round = 0; round < n - 1; ++round) {
            int pp = 0, qq = 0;
            if (act) {
                if (slot == 0) {
                    qq = 1 + (round % (n - 1));
                } else {
                    pp = 1 + (round + slot) % (n - 1);
                    qq = 1 + (round - slot + (n - 1)) % (n - 1);
                }
            }
            const int p = pp < qq ? pp : qq, q = pp < qq ? qq : pp;
            double2 *ap = reinterpret_cast<double2 *>(A + p * KP + 4 * l16);
            double2 *aq = reinterpret_cast<double2 *>(A + q * KP + 4 * l16);
            double2 *vp = reinterpret_cast<double2 *>(Vs + p * KP + 4 * l16);
            double2 *vq = reinterpret_cast<double2 *>(Vs + q * KP + 4 * l16);
            const double2 x0 = ap[0], x1 = ap[1], y0 = aq[0], y1 = aq[1];
            const double2 u0 = vp[0], u1 = vp[1], w0 = vq[0], w1 = vq[1];
            double al = x0.x * x0.x, be = y0.x * y0.x, ga = x0.x * y0.x;
            al = fma(x0.y, x0.y, al);
            be = fma(y0.y, y0.y, be);
            ga = fma(x0.y, y0.y, ga);
            al = fma(x1.x, x1.x, al);
            be = fma(y1.x, y1.x, be);
            ga = fma(x1.x, y1.x, ga);
            al = fma(x1.y, x1.y, al);
            be = fma(y1.y, y1.y, be);
            ga = fma(x1.y, y1.y, ga);
            al = row16_sum_dpp(al);
            be = row16_sum_dpp(be);
            ga = row16_sum_dpp(ga);
            if (act && ga != 0.0 && ga * ga > 1e-20 * (al * be)) {
                const double d = be - al, g2 = 2.0 * ga;
                const double hh = d * d + g2 * g2;
                const double h = hh * rsq_nr(hh);
                const double t = (d >= 0.0 ? g2 : -g2) / (fabs(d) + h);
                const double cs = rsq_nr(1.0 + t * t), sn = cs * t;
                ap[0] = make_double2(cs * x0.x - sn * y0.x, cs * x0.y - sn * y0.y);
                ap[1] = make_double2(cs * x1.x - sn * y1.x, cs * x1.y - sn * y1.y);
                aq[0] = make_double2(sn * x0.x + cs * y0.x, sn * x0.y + cs * y0.y);
                aq[1] = make_double2(sn * x1.x + cs * y1.x, sn * x1.y + cs * y1.y);
                vp[0] = make_double2(cs * u0.x - sn * w0.x, cs * u0.y - sn * w0.y);
                vp[1] = make_double2(cs * u1.x - sn * w1.x, cs * u1.y - sn * w1.y);
                vq[0] = make_double2(sn * u0.x + cs * w0.x, sn * u0.y + cs * w0.y);
                vq[1] = make_double2(sn * u1.x + cs * w1.x, sn * u1.y + cs * w1.y);
                changed[sweep % 3] = 1;
                if (ga * ga > 1e-10 * (al * be)) large[sweep % 3] = 1;
            }
            __syncthreads();
        }
        if (!changed[sweep % 3] || !large[sweep % 3]) break;
    }
    if (sweep == 60 && tid == 0) atomicOr(status, 2);
    for (int c = slot; c < KP; c += 32) {
        const double2 *ac = reinterpret_cast<const double2 *>(A + c * KP + 4 * l16);
        const double2 a0 = ac[0], a1 = ac[1];
        double s = a0.x * a0.x;
        s = fma(a0.y, a0.y, s);
        s = fma(a1.x, a1.x, s);
        s = fma(a1.y, a1.y, s);
        s = row16_sum_dpp(s);
        if (l16 == 0) nrm[c] = c < k ? sqrt(s) : -1.0;
    }
    __syncthreads();
    for (int c = tid; c < KP; c += blockDim.x) {
        int rank = 0;
#pragma unroll 4
        for (int d = 0; d < KP; ++d)
            if (nrm[d] > nrm[c] || (nrm[d] == nrm[c] && d < c)) ++rank;
        order[rank] = c;
    }
    __syncthreads();
    for (int t = tid; t < KP * KP; t += blockDim.x) {
        const int i = t / KP, jj = t % KP;
        const int c = order[jj];
        const double sg = nrm[c] > 0.0 ? nrm[c] : 0.0;
        U[t] = (jj < k && sg > 0.0) ? A[c * KP + i] / sg : 0.0;
        V[t] = jj < k ? Vs[c * KP + i] : 0.0;
    }
    for (int jj = tid; jj < KP; jj += blockDim.x)
        sigma[jj] = jj < k ? (nrm[order[jj]] > 0 ? nrm[order[jj]] : 0.0) : 0.0;
}

// Column statistics on the tile-CSC (values are X~): block (tile range,
// 16 waves x 8 genes); per tile the chunk's entries are staged by all threads
// into LDS, then lane l of a gene's wave takes
// run entries l, l+64, ... into register partial sums (f64) kept across the
// tile range; one wave reduction per gene at the end (deterministic).
__global__ void __launch_bounds__(1024) k_col_stats4(const int64_t *__restrict__ indptr,
                                                     const uint32_t *__restrict__ tptr,
                                                     const uint8_t *__restrict__ trow,
                                                     const float *__restrict__ tval, int nch,
                                                     int tiles_per_tr, int n_tiles, int64_t n_obs,
                                                     int64_t n_vars, double *__restrict__ psum,
                                                     double *__restrict__ psq, uint32_t *__restrict__ pcnt) {
    constexpr int GWS = 8, NW = 16, CGS = NW * GWS, ENT = 16384;
    __shared__ float xs[ENT];
    __shared__ uint32_t bnd[CGS + 1];
    const int tr = blockIdx.x / nch, gc = blockIdx.x % nch;
    const int wave = threadIdx.x / kWave, lane = threadIdx.x & 63;
    const int64_t cb = static_cast<int64_t>(gc) * CGS;
    const int64_t cw = cb + wave * GWS;
    const int64_t cend = (cb + CGS) < n_vars ? (cb + CGS) : n_vars;
    double s[GWS], q[GWS];
    uint32_t n[GWS];
#pragma unroll
    for (int gi = 0; gi < GWS; ++gi) { s[gi] = 0.0; q[gi] = 0.0; n[gi] = 0u; }
    const int t_begin = tr * tiles_per_tr;
    const int t_end = (t_begin + tiles_per_tr) < n_tiles ? (t_begin + tiles_per_tr) : n_tiles;
    for (int t = t_begin; t < t_end; ++t) {
        const int64_t r0 = static_cast<int64_t>(t) * kTileRows;
        const int64_t tb = indptr[r0];
        const uint32_t *tp = tptr + static_cast<int64_t>(t) * (n_vars + 1);
        __syncthreads();
        for (int i = threadIdx.x; i <= CGS; i += blockDim.x) bnd[i] = (cb + i <= cend) ? tp[cb + i] : 0u;
        __syncthreads();
        const uint32_t e_lo = (cb < n_vars) ? bnd[0] : 0u, e_hi = (cb < n_vars) ? bnd[cend - cb] : 0u;
        for (uint32_t w0 = e_lo; w0 < e_hi; w0 += ENT) {
            const uint32_t w1 = (w0 + ENT) < e_hi ? (w0 + ENT) : e_hi;
            for (uint32_t i = w0 + threadIdx.x; i < w1; i += blockDim.x)
                xs[i - w0] = tval[tb + i];
            __syncthreads();
#pragma unroll
            for (int gi = 0; gi < GWS; ++gi) {
                const int64_t c = cw + gi;
                if (c < cend) {
                    uint32_t a = bnd[c - cb], b = bnd[c - cb + 1];
                    if (w0 == e_lo) n[gi] += b - a;
                    a = a > w0 ? a : w0;
                    b = b < w1 ? b : w1;
                    for (uint32_t i = a + lane; i < b; i += kWave) {
                        const double x = static_cast<double>(xs[i - w0]);
                        s[gi] += x;
                        q[gi] += x * x;
                    }
                }
            }
            __syncthreads();
        }
    }
#pragma unroll
    for (int gi = 0; gi < GWS; ++gi) {
        const double ts = wave_sum(s[gi]), tq = wave_sum(q[gi]);
        const int64_t c = cw + gi;
        if (lane == 0 && c < cend) {
            const int64_t o = static_cast<int64_t>(tr) * n_vars + c;
            psum[o] = ts;
            psq[o] = tq;
            pcnt[o] = n[gi];
        }
    }
}

}  // namespace scxpca




