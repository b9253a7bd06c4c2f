// bf16 MFMA SpMMs for the BF16 sketch form (gfx950).
//
// The sketch passes of the BF16 form multiply X~ rounded to bf16 (the
// tile-CSC's txbf copy) by the f32 panels. Here the sparse operand is
// densified, a 32-gene block at a time, in LDS and multiplied on the matrix
// cores with v_mfma_f32_32x32x16_bf16 (f32 accumulation). The panel is split
// into two bf16 terms, hi = bf16(p) and lo = bf16(p - hi), and both are
// multiplied (two MFMAs per product), so the panel keeps ~16 significant bits
// and the only rounding of the form is the one of X~ to bf16.
//
// Lane layouts of v_mfma_f32_32x32x16_bf16 (checked on the GPU): lane l holds
// A[i = l % 32][k = 8 (l / 32) + e] and B[k = 8 (l / 32) + e][j = l % 32],
// e = 0..7; accumulator register r holds D[8 (r / 4) + 4 (l / 32) + r % 4][l % 32].
//
// Panels in the B-operand layout ("hilo-packed"): a row-major f32 matrix
// (K x KP) becomes two bf16 arrays in which element (k, c) sits at
// ((k / 8) * KP + c) * 8 + k % 8, so one lane's operand (8 consecutive k of
// one column) is one 16-byte load.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace scxpca {

typedef __bf16 bf16x8_t __attribute__((ext_vector_type(8)));
typedef float f32x16_t __attribute__((ext_vector_type(16)));

__device__ __forceinline__ f32x16_t mfma_bf16(const uint4 &a, const uint4 &b, const f32x16_t &c) {
    return __builtin_amdgcn_mfma_f32_32x32x16_bf16(__builtin_bit_cast(bf16x8_t, a), __builtin_bit_cast(bf16x8_t, b), c,
                                                   0, 0, 0);
}

// hi / lo split of a row-major f32 matrix (rows x KP; rows up to kpad are zero)
// into the hilo-packed layout. One thread per (8-row group, column).
template <int KP>
__global__ void __launch_bounds__(256) k_pack_hilo(const float *__restrict__ A, int64_t rows, int64_t kpad,
                                                   uint16_t *__restrict__ hi, uint16_t *__restrict__ lo) {
    const int64_t t = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (t >= kpad / 8 * KP) return;
    const int64_t kg = t / KP;
    const int c = static_cast<int>(t % KP);
    uint32_t h[4], l[4];
#pragma unroll
    for (int e = 0; e < 8; e += 2) {
        float a[2];
#pragma unroll
        for (int u = 0; u < 2; ++u) {
            const int64_t k = 8 * kg + e + u;
            a[u] = k < rows ? A[k * KP + c] : 0.f;
        }
        uint16_t hb[2], lb[2];
#pragma unroll
        for (int u = 0; u < 2; ++u) {
            hb[u] = f32_to_bf16_rne(a[u]);
            lb[u] = f32_to_bf16_rne(a[u] - bf16_to_f32(hb[u]));
        }
        h[e / 2] = static_cast<uint32_t>(hb[0]) | (static_cast<uint32_t>(hb[1]) << 16);
        l[e / 2] = static_cast<uint32_t>(lb[0]) | (static_cast<uint32_t>(lb[1]) << 16);
    }
    reinterpret_cast<uint4 *>(hi)[t] = make_uint4(h[0], h[1], h[2], h[3]);
    reinterpret_cast<uint4 *>(lo)[t] = make_uint4(l[0], l[1], l[2], l[3]);
}

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

constexpr int kTG = 64;                   // genes per wave of the transposed kernel (two 32-row A blocks)
constexpr int kTSlabLd = 128 + 8;         // transposed slab row (bf16): a 128-row half-tile, padded
constexpr int kFG = 64;                   // genes per chunk of the forward kernel
constexpr int kFSlabLd = kFG + 8;         // forward slab row (bf16), padded

// [s0, s1) of a tile's 64-gene range (tile-relative entries), from run
// starts held by lanes 0 and 1
__device__ __forceinline__ void range_of(uint32_t starts, uint32_t &s0, uint32_t &s1) {
    s0 = static_cast<uint32_t>(__builtin_amdgcn_readlane(static_cast<int>(starts), 0));
    s1 = static_cast<uint32_t>(__builtin_amdgcn_readlane(static_cast<int>(starts), 1));
}

// ---------------------------------------------------------------------------
// Transposed: part[tr][gene][:] = sum over the tiles of range tr of
// bf16(X~)[tile, gene]^T (Q_hi + Q_lo)[tile, :], for the 64 columns of
// column half ch. Block: NWT waves, wave w owns the 64 genes g0 = 64 (NWT gg
// + w) .. g0 + 63, one contiguous tile-CSC range per tile. Per 128-row
// half-tile h the block stages the half-tile's packed Q (hi and lo, 64
// columns: 32 KB) into one of two LDS buffers (one barrier per half-tile);
// each wave zero-fills its private slab [64 genes][128 rows] (bf16), writes
// its entries of rows in half h (packed entries carry gene and row), and runs
// 8 k-steps x 2 gene blocks x 2 column blocks x {hi, lo} MFMAs (per k-step
// 2 A and 4 B fragments feed 8 MFMAs). A tile's first NBP entry batches are
// loaded while the previous tile is multiplied (two register sets, the tile
// loop unrolled by two), its range two tiles ahead. Blocks are XCD-aware:
// blockIdx = slot * 8 + xcd, the gene groups and column halves of one tile
// range on one XCD (their Q half-tiles shared in L2). Grid: nrb * ngg * ncolh
// blocks, nrb a multiple of 8.
// ---------------------------------------------------------------------------
template <int NWT, int NBP>
__global__ void __launch_bounds__(NWT * 64) k_spmm_t_mfma(const int64_t *__restrict__ indptr,
                                                       const uint32_t *__restrict__ tptr,
                                                       const uint32_t *__restrict__ tpk,
                                                       const uint16_t *__restrict__ qhi,
                                                       const uint16_t *__restrict__ qlo, int kp, int ngg, int ncolh,
                                                       int nrb, int tiles_per_tr, int64_t n_vars, int n_tiles,
                                                       float *__restrict__ part) {
    constexpr int QU4 = 16 * 64;  // uint4 per half-tile of one term (16 k-groups x 64 columns)
    __shared__ uint4 qs[2][2][QU4];                     // [buffer][hi, lo]
    __shared__ uint4 slab[NWT][kTG * kTSlabLd / 8];     // per-wave [gene][row] bf16
    const int tid = threadIdx.x, lane = tid & 63;
    const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
    const int xcd = blockIdx.x & 7, slot = blockIdx.x >> 3;
    const int per = ngg * ncolh;
    const int tr = (slot / per) * 8 + xcd;
    const int gg = (slot % per) / ncolh, ch = (slot % per) % ncolh;
    const int64_t g0 = (static_cast<int64_t>(gg) * NWT + wave) * kTG;
    const int t_begin = tr * tiles_per_tr;
    const int t_end = (tr < nrb) ? ((t_begin + tiles_per_tr) < n_tiles ? (t_begin + tiles_per_tr) : n_tiles) : t_begin;
    const int64_t nv1 = n_vars + 1;
    // lanes 0 and 1 read the wave's range bounds: run starts of g0 and g0 + 64
    const int64_t gl = g0 + (lane == 1 ? kTG : 0);
    const int64_t gslot = gl < n_vars ? gl : n_vars;
    uint16_t *my = reinterpret_cast<uint16_t *>(slab[wave]);
    f32x16_t acc[2][2];
#pragma unroll
    for (int mb = 0; mb < 2; ++mb)
#pragma unroll
        for (int nb = 0; nb < 2; ++nb)
#pragma unroll
            for (int r = 0; r < 16; ++r) acc[mb][nb][r] = 0.f;

    auto q_src = [&](int t, int h, int i, const uint16_t *base) {
        // uint4 i of the half-tile (k-group i / 64, column i % 64) of column half ch
        const int64_t kg = static_cast<int64_t>(t) * 32 + h * 16 + i / 64;
        return reinterpret_cast<const uint4 *>(base)[kg * kp + ch * 64 + (i % 64)];
    };
    constexpr int QPT = QU4 / (NWT * 64);  // uint4 per thread per term
    static_assert(QPT * NWT * 64 == QU4, "whole uint4s per thread");
    uint4 qh[QPT], ql[QPT];
    auto load_q = [&](int t, int h) {
#pragma unroll
        for (int j = 0; j < QPT; ++j) {
            qh[j] = q_src(t, h, tid + j * NWT * 64, qhi);
            ql[j] = q_src(t, h, tid + j * NWT * 64, qlo);
        }
    };
    auto store_q = [&](int buf) {
#pragma unroll
        for (int j = 0; j < QPT; ++j) {
            qs[buf][0][tid + j * NWT * 64] = qh[j];
            qs[buf][1][tid + j * NWT * 64] = ql[j];
        }
    };
    auto starts_of = [&](int t) -> uint32_t {
        return (t < t_end && lane < 2) ? tptr[static_cast<int64_t>(t) * nv1 + gslot] : 0u;
    };
    auto tile_base = [&](int t) -> int64_t { return t < t_end ? indptr[static_cast<int64_t>(t) * 256] : 0; };
    struct Range { uint32_t s0, s1; int64_t tb; };
    auto issue = [&](const Range &r, uint32_t (&raw)[NBP]) {
#pragma unroll
        for (int b = 0; b < NBP; ++b) {
            const uint32_t p = r.s0 + 64 * b + static_cast<uint32_t>(lane);
            raw[b] = tpk[r.tb + (p < r.s1 ? p : r.s0)];
        }
    };
    auto put = [&](uint32_t e, bool ok, int h) {
        if (ok && ((e >> 23) & 1u) == static_cast<uint32_t>(h))
            my[((e >> 24) & 63u) * kTSlabLd + ((e >> 16) & 127u)] = static_cast<uint16_t>(e);
    };
    // half h of a tile's entries into the (zero-filled) slab
    auto fill = [&](int h, const Range &r, const uint32_t (&raw)[NBP]) {
#pragma unroll
        for (int b = 0; b < NBP; ++b) put(raw[b], r.s0 + 64 * b + static_cast<uint32_t>(lane) < r.s1, h);
        for (uint32_t p0 = r.s0 + 64 * NBP; p0 < r.s1; p0 += 64) {  // dense ranges: direct reads
            const uint32_t p = p0 + static_cast<uint32_t>(lane);
            put(tpk[r.tb + (p < r.s1 ? p : r.s0)], p < r.s1, h);
        }
    };
    int buf = 0;
    // One tile: cur = its entries (loaded), rC its range; nxt <- tile t + 1's
    // entries (range rN from stN, loaded); stL <- tile t + 2's run starts
    auto tile = [&](int t, uint32_t (&cur)[NBP], uint32_t (&nxt)[NBP], Range &rC, uint32_t &stN, uint32_t &stL) {
        if (t >= t_end) return;
        Range rN{0u, 0u, tile_base(t + 1)};
        range_of(stN, rN.s0, rN.s1);
        if (t + 1 < t_end) issue(rN, nxt);
        stL = starts_of(t + 2);
        for (int h = 0; h < 2; ++h) {
            const bool more = (h == 0) || (t + 1 < t_end);
            if (more) load_q(h == 0 ? t : t + 1, h ^ 1);
            for (int i = lane; i < kTG * kTSlabLd / 8; i += 64) slab[wave][i] = make_uint4(0u, 0u, 0u, 0u);
            wave_lds_sync();
            fill(h, rC, cur);
            __syncthreads();  // Q buffer `buf` complete; every wave's slab written
            const uint4 *qsh = qs[buf][0], *qsl = qs[buf][1];
#pragma unroll
            for (int s = 0; s < 8; ++s) {
                uint4 a[2];
#pragma unroll
                for (int mb = 0; mb < 2; ++mb)
                    a[mb] = *reinterpret_cast<const uint4 *>(my + (32 * mb + (lane & 31)) * kTSlabLd + 16 * s +
                                                             8 * (lane >> 5));
#pragma unroll
                for (int nb = 0; nb < 2; ++nb) {
                    const int bi = (2 * s + (lane >> 5)) * 64 + nb * 32 + (lane & 31);
                    const uint4 bh = qsh[bi], bl = qsl[bi];
#pragma unroll
                    for (int mb = 0; mb < 2; ++mb) {
                        acc[mb][nb] = mfma_bf16(a[mb], bh, acc[mb][nb]);
                        acc[mb][nb] = mfma_bf16(a[mb], bl, acc[mb][nb]);
                    }
                }
            }
            if (more) store_q(buf ^ 1);  // buffer buf ^ 1 was last read before this half's barrier
            buf ^= 1;
        }
        rC = rN;
    };
    uint32_t rawA[NBP], rawB[NBP];
    Range rC{0u, 0u, 0};
    uint32_t stX = 0u, stY = 0u;
    if (t_begin < t_end) {
        const uint32_t st0 = starts_of(t_begin);
        rC.tb = tile_base(t_begin);
        range_of(st0, rC.s0, rC.s1);
        issue(rC, rawA);
        load_q(t_begin, 0);
        store_q(0);
        stX = starts_of(t_begin + 1);
    }
    for (int t = t_begin; t < t_end; t += 2) {
        tile(t, rawA, rawB, rC, stX, stY);      // stX: tile t + 1's starts; tile t + 2's into stY
        tile(t + 1, rawB, rawA, rC, stY, stX);  // stY: tile t + 2's starts; tile t + 3's into stX
    }
    if (tr >= nrb) return;
#pragma unroll
    for (int mb = 0; mb < 2; ++mb)
#pragma unroll
        for (int nb = 0; nb < 2; ++nb)
#pragma unroll
            for (int r = 0; r < 16; ++r) {
                const int64_t gene = g0 + 32 * mb + 8 * (r / 4) + 4 * (lane >> 5) + (r % 4);
                if (gene < n_vars)
                    part[(static_cast<int64_t>(tr) * n_vars + gene) * kp + ch * 64 + nb * 32 + (lane & 31)] =
                        acc[mb][nb][r];
            }
}

// ---------------------------------------------------------------------------
// Forward: Y[rows of the block, column half ch] = bf16(X~) (M_hi + M_lo)
// - mc (+ rowb[r] cs[c], the PFlog baseline term). Block: 8 waves on two
// 256-row tiles; wave w owns the 64 rows 64 (w % 4) .. + 63 of tile w / 4.
// Per 64-gene chunk J every wave of a tile reads the tile's chunk range (one
// contiguous tile-CSC range) and writes the entries of its own rows into its
// private slab [64 rows][64 genes] (bf16), so slabs need no block
// synchronization; the chunk's packed M rows (hi and lo, 64 columns: 16 KB)
// are block-shared, in two LDS buffers (one barrier per chunk). Per chunk
// each wave runs 4 k-steps x 2 row blocks x 2 column blocks x {hi, lo}
// MFMAs, then zero-fills its slab. A chunk's first NBP entry batches are
// loaded while the previous chunk is multiplied (two register sets, the
// chunk loop unrolled by two), its range two chunks ahead.
// ---------------------------------------------------------------------------
template <bool OFFSET, int NBP>
__global__ void __launch_bounds__(512) k_spmm_fwd_mfma(const int64_t *__restrict__ indptr,
                                                     const uint32_t *__restrict__ tptr,
                                                     const uint32_t *__restrict__ tpk,
                                                     const uint16_t *__restrict__ mhi,
                                                     const uint16_t *__restrict__ mlo, int kp, int ncolh,
                                                     int64_t n_obs, int64_t n_vars, int n_tiles,
                                                     const float *__restrict__ mc, const double *__restrict__ rowb,
                                                     const float *__restrict__ cs, float *__restrict__ Y) {
    constexpr int SLAB_U4 = 64 * kFSlabLd / 8;  // uint4 per wave slab
    constexpr int MU4 = 8 * 64;                 // uint4 per chunk of one term (8 k-groups x 64 columns)
    __shared__ uint4 slab[8][SLAB_U4];
    __shared__ uint4 ms[2][2][MU4];
    const int tid = threadIdx.x, lane = tid & 63;
    const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
    const int pair = static_cast<int>(blockIdx.x) / ncolh, ch = static_cast<int>(blockIdx.x) % ncolh;
    const int t0 = 2 * pair;
    const int u = wave >> 2;                 // this wave's tile
    const uint32_t qw = static_cast<uint32_t>(wave & 3);  // its 64-row quarter
    const bool has_tile = t0 + u < n_tiles;
    const int64_t nv1 = n_vars + 1;
    const int nch = static_cast<int>((n_vars + kFG - 1) / kFG);
    const int64_t tbu = has_tile ? indptr[static_cast<int64_t>(t0 + u) * 256] : 0;
    uint16_t *my = reinterpret_cast<uint16_t *>(slab[wave]);
    f32x16_t acc[2][2];
#pragma unroll
    for (int mb = 0; mb < 2; ++mb)
#pragma unroll
        for (int nb = 0; nb < 2; ++nb)
#pragma unroll
            for (int r = 0; r < 16; ++r) acc[mb][nb][r] = 0.f;
    // run starts of chunk J's bounds (lanes 0 and 1); 0 past the chunks
    auto starts_of = [&](int J) -> uint32_t {
        if (J >= nch || !has_tile || lane > 1) return 0u;
        const int64_t g = static_cast<int64_t>(J) * kFG + (lane == 1 ? kFG : 0);
        return tptr[static_cast<int64_t>(t0 + u) * nv1 + (g < n_vars ? g : n_vars)];
    };
    auto issue = [&](uint32_t s0, uint32_t s1, uint32_t (&raw)[NBP]) {
#pragma unroll
        for (int b = 0; b < NBP; ++b) {
            const uint32_t p = s0 + 64 * b + static_cast<uint32_t>(lane);
            raw[b] = tpk[tbu + (p < s1 ? p : s0)];
        }
    };
    auto put = [&](uint32_t e, bool ok) {
        const uint32_t row = (e >> 16) & 255u;
        if (ok && (row >> 6) == qw) my[(row & 63u) * kFSlabLd + ((e >> 24) & 63u)] = static_cast<uint16_t>(e);
    };
    auto scatter = [&](uint32_t s0, uint32_t s1, const uint32_t (&raw)[NBP]) {
#pragma unroll
        for (int b = 0; b < NBP; ++b) put(raw[b], s0 + 64 * b + static_cast<uint32_t>(lane) < s1);
        for (uint32_t p0 = s0 + 64 * NBP; p0 < s1; p0 += 64) {  // dense chunks: direct reads
            const uint32_t p = p0 + static_cast<uint32_t>(lane);
            put(tpk[tbu + (p < s1 ? p : s0)], p < s1);
        }
    };
    uint4 mh, ml;  // one uint4 of each term per thread
    auto load_m = [&](int J) {
        const int64_t kg = static_cast<int64_t>(J) * 8 + tid / 64;
        mh = reinterpret_cast<const uint4 *>(mhi)[kg * kp + ch * 64 + (tid % 64)];
        ml = reinterpret_cast<const uint4 *>(mlo)[kg * kp + ch * 64 + (tid % 64)];
    };
    auto store_m = [&](int b) {
        ms[b][0][tid] = mh;
        ms[b][1][tid] = ml;
    };
    for (int i = lane; i < SLAB_U4; i += 64) slab[wave][i] = make_uint4(0u, 0u, 0u, 0u);
    // One chunk: cur = its entries (loaded), [c0, c1) its range; nxt <- chunk
    // J + 1's entries (range from stN, loaded); stL <- chunk J + 2's bounds
    uint32_t c0 = 0u, c1 = 0u;
    auto chunk = [&](int J, uint32_t (&cur)[NBP], uint32_t (&nxt)[NBP], uint32_t &stN, uint32_t &stL) {
        if (J >= nch) return;
        const bool next = J + 1 < nch;
        if (next) load_m(J + 1);
        uint32_t n0 = 0u, n1 = 0u;
        range_of(stN, n0, n1);
        if (next && has_tile) issue(n0, n1, nxt);
        stL = starts_of(J + 2);
        __syncthreads();  // chunk J's M complete in ms[J & 1]; ms[(J + 1) & 1] no longer read
        if (has_tile) scatter(c0, c1, cur);
        wave_lds_sync();
        const int b = J & 1;
#pragma unroll
        for (int s = 0; s < 4; ++s) {
            uint4 a[2];
#pragma unroll
            for (int mb = 0; mb < 2; ++mb)
                a[mb] = *reinterpret_cast<const uint4 *>(my + (32 * mb + (lane & 31)) * kFSlabLd + 16 * s +
                                                         8 * (lane >> 5));
#pragma unroll
            for (int nb = 0; nb < 2; ++nb) {
                const int bi = (2 * s + (lane >> 5)) * 64 + nb * 32 + (lane & 31);
                const uint4 bh = ms[b][0][bi], bl = ms[b][1][bi];
#pragma unroll
                for (int mb = 0; mb < 2; ++mb) {
                    acc[mb][nb] = mfma_bf16(a[mb], bh, acc[mb][nb]);
                    acc[mb][nb] = mfma_bf16(a[mb], bl, acc[mb][nb]);
                }
            }
        }
        // the slab's reads above complete before these writes (one wave, in order)
        for (int i = lane; i < SLAB_U4; i += 64) slab[wave][i] = make_uint4(0u, 0u, 0u, 0u);
        if (next) store_m((J + 1) & 1);
        c0 = n0;
        c1 = n1;
    };
    uint32_t rawA[NBP], rawB[NBP];
    uint32_t stX = 0u, stY = 0u;
    if (nch > 0) {
        const uint32_t st0 = starts_of(0);
        range_of(st0, c0, c1);
        if (has_tile) issue(c0, c1, rawA);
        load_m(0);
        store_m(0);
        stX = starts_of(1);
    }
    for (int J = 0; J < nch; J += 2) {
        chunk(J, rawA, rawB, stX, stY);      // stX: chunk J + 1's bounds; chunk J + 2's into stY
        chunk(J + 1, rawB, rawA, stY, stX);  // stY: chunk J + 2's bounds; chunk J + 3's into stX
    }
    const int64_t r0 = static_cast<int64_t>(t0 + u) * 256 + 64 * static_cast<int64_t>(qw);
#pragma unroll
    for (int mb = 0; mb < 2; ++mb)
#pragma unroll
        for (int r = 0; r < 16; ++r) {
            const int64_t row = r0 + 32 * mb + 8 * (r / 4) + 4 * (lane >> 5) + (r % 4);
            if (!has_tile || row >= n_obs) continue;
            const float br = OFFSET ? static_cast<float>(rowb[row]) : 0.f;
#pragma unroll
            for (int nb = 0; nb < 2; ++nb) {
                const int col = ch * 64 + nb * 32 + (lane & 31);
                float o = acc[mb][nb][r] - (mc ? mc[col] : 0.f);
                if constexpr (OFFSET) o += br * cs[col];
                Y[row * kp + col] = o;
            }
        }
}

}  // namespace scxpca
