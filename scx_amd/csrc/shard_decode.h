// On-GPU decode of SCX CSR shards (C ABI in include/scxshard.h).
//
// Included by scxpca.hip after its error/device helpers (Error, HIPC,
// guarded, DevBuf, DeviceGuard, PinnedBuf).
//
// Design (one call decodes any number of shards into one row-stacked CSR):
//  * host, in parallel over row groups (a framed shard's block-index entries,
//    or the whole of an unframed shard): header and block-index checks; the
//    group's indptr (Delta-Golomb, raw, or a zstd / LZ4 frame -- a few bytes per
//    row) into the global i64 indptr; the group's index and value payloads into
//    one pinned staging buffer (the encoded bytes for Scx1 and None, the
//    decompressed byte planes for the frame codecs); and the work lists the
//    kernels need -- one entry per non-empty FOR-BP row (byte offset, frame
//    min, frame bits, layout) and one per Rice block (bit offset, k). Finding
//    those offsets walks headers and skips payloads: rows cost O(1) each, Rice
//    blocks with k = 0 are skipped by counting zero bits 64 at a time.
//  * one host->device copy of the staging buffer and the work lists;
//  * device: k_forbp_rows (a wavefront per row: every lane extracts one delta,
//    a wave prefix sum rebuilds the indices, coalesced i32 stores with the
//    n_minor / 2^31 bound check), k_rice_blocks (a lane per 256-value block,
//    64-bit windows and a count-trailing-ones per unary code, f32 stores),
//    k_plane_sums / k_plane_carry / k_plane_emit (byte un-delta as a prefix sum
//    of packed-byte words -- four mod-256 lanes in one u32 -- fused with the
//    un-shuffle and the widening to i32 / f32).
#pragma once

#include <dlfcn.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <cstring>
#include <functional>
#include <string>
#include <thread>
#include <vector>

#include "../../include/scxshard.h"

namespace shardgpu {

constexpr int kHdr = 76;       // SHARD_HEADER_SIZE (scx-format/src/shard.rs:10)
constexpr int kBie = 22;       // BLOCK_INDEX_ENTRY_SIZE (scx-format/src/shard.rs:33)
constexpr int kBVal = 256;     // Rice block (scx-codec/src/rice.rs:15)
constexpr int kBIdx = 128;     // FOR-BP rows per block (scx-codec/src/forbp.rs:12)
constexpr int kSimd = 128;     // BitPacker4x threshold (scx-codec/src/forbp.rs:19)
constexpr int kPlaneChunk = 2048;  // elements per workgroup of the plane kernels
constexpr int kStagePad = 64;      // slack after every staged payload (window reads)

enum Codec { C_NONE = 0, C_SCX1 = 1, C_ZSTD = 2, C_LZ4 = 3, C_PCODEC = 4, C_SHUF = 5 };
enum Venc { V_U8 = 0, V_U16 = 1, V_U32 = 2, V_F32 = 3, V_F16 = 4 };

// ---------------------------------------------------------------------------
// Device work lists
// ---------------------------------------------------------------------------
struct RowJob {        // one non-empty FOR-BP row
    uint64_t src;      // byte offset of the packed deltas in the staging buffer
    uint64_t out;      // first output element
    uint32_t nnz;
    uint32_t fmin;
    uint32_t meta;     // frame_bits | (BitPacker4x layout) << 8
    uint32_t pad;
};
struct RiceJob {       // one Rice block
    uint64_t bit;      // bit offset of the first code in the staging buffer
    uint64_t out;
    uint32_t nk;       // values | k << 16
    uint32_t nbits;    // bits from `bit` to the end of the block's last code
};
constexpr int kRiceLdsWords = 4096;  // per wave: 16 KB of staged block bytes
struct PlaneJob {      // one index or value sub-stream held as bytes
    uint64_t src;      // staging byte offset
    uint64_t out;      // first output element
    uint64_t n;        // elements
    uint32_t width;    // bytes per element (1, 2, 4)
    uint32_t flags;    // kPlaneMajor | kDelta | kIndex | value encoding << 8
    uint64_t chunk0;   // first chunk of this job in the chunk table
};
constexpr uint32_t kPlaneMajor = 1, kDelta = 2, kIndex = 4;
struct ChunkJob {
    uint32_t job;
    uint32_t pad;
    uint64_t e0;       // first element of the chunk within its job
};
enum : uint32_t { ERR_BOUND = 1, ERR_I32 = 2, ERR_U32 = 4 };

// 64 bits of the staging buffer starting at bit `bit` (the low 33 or more are
// the requested bits). The buffer is 16-byte aligned and padded.
__device__ __forceinline__ uint64_t window(const uint32_t *__restrict__ w, uint64_t bit) {
    const uint64_t a = bit >> 5;
    const uint64_t v = (uint64_t)w[a] | ((uint64_t)w[a + 1] << 32);
    return v >> (bit & 31);
}

// FOR-BP rows: a wavefront per row, four rows per workgroup. Each lane takes
// four consecutive entries of a 256-entry slice, extracts their deltas, and a
// wave prefix sum of the lane sums plus the running carry (frame_min + earlier
// slices) rebuilds the indices.
// (forbp_decode_kernel / forbp_decode_bp4x_kernel semantics,
// scx-gpu/kernels/forbp_decode.cu:13-64, forbp_decode_bp4x.cu:41-119; the
// u32-overflow and index-bound checks of scx-codec's decoder and
// u32_vec_to_i32_bounded, scx-codec/src/forbp.rs:478-483, dispatch.rs:250.)
__device__ __forceinline__ uint32_t forbp_delta(const uint32_t *__restrict__ stage, const RowJob &r, uint32_t i,
                                                uint32_t fb, uint64_t mask, uint32_t full, uint64_t tail_bit) {
    if (i < full) {
        // value p of a 128-chunk: lane p & 3, slot p >> 2 of that lane's
        // little-endian u32 stream at 16-byte stride
        const uint32_t c = i / kSimd, p = i % kSimd, l = p & 3u, slot = p >> 2;
        const uint32_t bitpos = slot * fb, w0 = bitpos >> 5, bin = bitpos & 31u;
        const uint64_t byte = r.src + (uint64_t)c * fb * 16u + w0 * 16u + l * 4u;
        uint64_t v = (uint32_t)window(stage, byte * 8u) >> bin;
        if (bin + fb > 32u) v |= (uint64_t)(uint32_t)window(stage, (byte + 16u) * 8u) << (32u - bin);
        return (uint32_t)(v & mask);
    }
    return (uint32_t)(window(stage, tail_bit + (uint64_t)(i - full) * fb) & mask);
}

__global__ __launch_bounds__(256) void k_forbp_rows(const uint32_t *__restrict__ stage,
                                                    const RowJob *__restrict__ jobs, int64_t n_jobs,
                                                    int32_t *__restrict__ out, uint32_t bound,
                                                    uint32_t *__restrict__ err) {
    __shared__ uint32_t xbuf[4][256];  // a wave's slice, for coalesced stores
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int64_t j = (int64_t)blockIdx.x * 4 + wave;
    if (j >= n_jobs) return;
    const RowJob r = jobs[j];
    const uint32_t fb = r.meta & 0xffu;
    const bool bp4x = (r.meta >> 8) & 1u;
    const uint64_t mask = fb >= 32 ? 0xffffffffull : ((1ull << fb) - 1);
    const uint32_t full = bp4x ? r.nnz / kSimd * kSimd : 0u;
    const uint64_t tail_bit = (r.src + (uint64_t)(full / kSimd) * fb * 16u) * 8u;
    uint64_t carry = r.fmin;
    uint32_t bad = 0;
    uint32_t *xb = xbuf[wave];
    for (uint32_t base = 0; base < r.nnz; base += 256) {
        const uint32_t i0 = base + (uint32_t)lane * 4u;
        uint32_t d[4];
        uint64_t sum = 0;
#pragma unroll
        for (int e = 0; e < 4; ++e) {
            d[e] = (i0 + e < r.nnz && fb) ? forbp_delta(stage, r, i0 + e, fb, mask, full, tail_bit) : 0u;
            sum += d[e];
        }
        uint64_t s = sum;
#pragma unroll
        for (int off = 1; off < 64; off <<= 1) {
            const uint64_t t = __shfl_up(s, off, 64);
            if (lane >= off) s += t;
        }
        uint64_t idx = carry + s - sum;
#pragma unroll
        for (int e = 0; e < 4; ++e) {
            const uint32_t i = i0 + e;
            idx += d[e];
            if (i < r.nnz) {
                if (idx > 0xffffffffull) bad |= ERR_U32;
                else if (idx > 0x7fffffffull) bad |= ERR_I32;
                else if (bound && idx >= bound) bad |= ERR_BOUND;
            }
            xb[lane * 4 + e] = (uint32_t)idx;
        }
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
#pragma unroll
        for (int e = 0; e < 4; ++e) {
            const uint32_t i = base + (uint32_t)e * 64u + (uint32_t)lane;
            if (i < r.nnz) out[r.out + i] = (int32_t)xb[e * 64 + lane];
        }
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        carry += __shfl(s, 63, 64);
    }
    if (bad) atomicOr(err, bad);
}

// Rice blocks: a lane per block of up to 256 values, 64 consecutive blocks
// per wave. The blocks of a wave are adjacent in the staging buffer, so the
// wave first copies their bytes into LDS with coalesced loads and every lane
// then walks its codes out of LDS (a serial chain of LDS reads instead of
// global-memory round trips); a span above the LDS budget is walked in place.
// The host walk already checked every code, so no read leaves the span.
// (rice_decode_kernel semantics, scx-gpu/kernels/rice_decode.cu:14-88:
// value = ((q << k) | r) + 1, then the u32 -> f32 conversion of
// decode_shard_scipy, scx-codec/src/dispatch.rs:252.)
__global__ __launch_bounds__(256) void k_rice_blocks(const uint32_t *__restrict__ stage,
                                                     const RiceJob *__restrict__ jobs, int64_t n_jobs,
                                                     float *__restrict__ out) {
    __shared__ uint32_t lds[4][kRiceLdsWords + 2];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int64_t j0 = ((int64_t)blockIdx.x * 4 + wave) * 64;
    const int64_t jn = j0 < n_jobs ? (n_jobs - j0 < 64 ? n_jobs : j0 + 64) : j0;
    const int64_t j = j0 + lane;
    RiceJob b{0, 0, 0, 0};
    if (j < jn) b = jobs[j];
    bool in_lds = false;
    uint64_t w0 = 0;
    if (jn > j0) {
        const uint64_t first = __shfl(b.bit, 0, 64);
        const int last = (int)(jn - j0 - 1);
        const uint64_t end = __shfl(b.bit + b.nbits, last, 64);
        w0 = first >> 5;
        const uint64_t nw = ((end + 31) >> 5) - w0 + 2;
        in_lds = nw <= (uint64_t)kRiceLdsWords;
        if (in_lds)
            for (uint64_t w = lane; w < nw; w += 64) lds[wave][w] = stage[w0 + w];
    }
    __syncthreads();
    if (j >= jn) return;
    const uint32_t n = b.nk & 0xffffu, k = b.nk >> 16;
    const uint32_t mask = (1u << k) - 1u;
    const uint32_t *src = in_lds ? lds[wave] : stage;
    uint64_t p = in_lds ? b.bit - w0 * 32 : b.bit;
    float *o = out + b.out;
    for (uint32_t i = 0; i < n; ++i) {
        uint32_t q = 0;
        for (;;) {
            const uint64_t w = window(src, p);
            const uint32_t valid = 64u - (uint32_t)(p & 31);
            const uint64_t inv = ~w;
            const uint32_t t = inv ? (uint32_t)__builtin_ctzll(inv) : 64u;
            if (t < valid) {
                q += t;
                p += t + 1;
                break;
            }
            q += valid;
            p += valid;
        }
        uint32_t r = 0;
        if (k) {
            r = (uint32_t)window(src, p) & mask;
            p += k;
        }
        o[i] = (float)(((q << k) | r) + 1u);
    }
}

// Rice blocks with k = 0, a wavefront per block: every code is a run of ones
// closed by a zero, so value i is the distance between the i-th and the
// (i-1)-th zero of the body. Lane l takes bits [32 l, 32 l + 32) of each
// 2,048-bit slice; a wave prefix sum of the zero counts numbers the codes, a
// wave max-scan of the last zero positions links the slices, the values go
// to LDS and leave in coalesced stores.
__global__ __launch_bounds__(256) void k_rice_k0(const uint32_t *__restrict__ stage,
                                                 const RiceJob *__restrict__ jobs, int64_t n_jobs,
                                                 float *__restrict__ out) {
    __shared__ float obuf[4][kBVal];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int64_t j = (int64_t)blockIdx.x * 4 + wave;
    if (j >= n_jobs) return;
    const RiceJob b = jobs[j];
    const uint32_t n = b.nk & 0xffffu, nbits = b.nbits;
    uint32_t base = 0;
    int32_t last = -1;  // bit of the previous code's closing zero
    for (uint32_t c0 = 0; c0 < nbits; c0 += 2048) {
        const uint32_t at = c0 + (uint32_t)lane * 32u;
        uint32_t z = 0;
        if (at < nbits) {
            z = ~(uint32_t)window(stage, b.bit + at);
            const uint32_t valid = nbits - at;
            if (valid < 32) z &= (1u << valid) - 1u;
        }
        const uint32_t cnt = (uint32_t)__builtin_popcount(z);
        uint32_t incl = cnt;
        int32_t mx = z ? (int32_t)(at + 31u - (uint32_t)__builtin_clz(z)) : -1;
#pragma unroll
        for (int off = 1; off < 64; off <<= 1) {
            const uint32_t t = __shfl_up(incl, off, 64);
            const int32_t m = __shfl_up(mx, off, 64);
            if (lane >= off) {
                incl += t;
                mx = m > mx ? m : mx;
            }
        }
        // exclusive forms: the codes before this lane, the last zero before it
        uint32_t idx = base + incl - cnt;
        int32_t prev = __shfl_up(mx, 1, 64);
        if (lane == 0) prev = -1;
        prev = prev > last ? prev : last;
        while (z) {
            const int32_t zp = (int32_t)(at + (uint32_t)__builtin_ctz(z));
            if (idx < (uint32_t)kBVal) obuf[wave][idx] = (float)(uint32_t)(zp - prev);
            prev = zp;
            ++idx;
            z &= z - 1u;
        }
        base += __shfl(incl, 63, 64);
        const int32_t lm = __shfl(mx, 63, 64);
        last = lm > last ? lm : last;
    }
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    for (uint32_t i = (uint32_t)lane; i < n; i += 64) out[b.out + i] = obuf[wave][i];
}

// Packed-byte add: four independent mod-256 sums in one u32.
__host__ __device__ __forceinline__ uint32_t padd(uint32_t a, uint32_t b) {
    return (((a & 0x00ff00ffu) + (b & 0x00ff00ffu)) & 0x00ff00ffu) |
           (((a & 0xff00ff00u) + (b & 0xff00ff00u)) & 0xff00ff00u);
}

// Element e of a job as a little-endian word (bytes j at plane j * n + e, or
// at e * width + j when the stream is element-major).
__device__ __forceinline__ uint32_t elem_word(const uint8_t *__restrict__ src, uint64_t n, uint32_t width,
                                              bool plane_major, uint64_t e) {
    uint32_t v = 0;
    for (uint32_t b = 0; b < width; ++b) {
        const uint64_t at = plane_major ? (uint64_t)b * n + e : e * width + b;
        v |= (uint32_t)src[at] << (8u * b);
    }
    return v;
}

__device__ __forceinline__ uint32_t block_scan_padd(uint32_t v, uint32_t *lds, uint32_t &total) {
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    uint32_t s = v;
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
        const uint32_t t = __shfl_up(s, off, 64);
        if (lane >= off) s = padd(s, t);
    }
    if (lane == 63) lds[wave] = s;
    __syncthreads();
    uint32_t before = 0;
    for (int w = 0; w < wave; ++w) before = padd(before, lds[w]);
    total = 0;
    for (int w = 0; w < 4; ++w) total = padd(total, lds[w]);
    __syncthreads();
    return padd(before, s);  // inclusive
}

// Per-chunk packed sums of the delta-coded streams.
__global__ __launch_bounds__(256) void k_plane_sums(const uint8_t *__restrict__ stage,
                                                    const PlaneJob *__restrict__ jobs,
                                                    const ChunkJob *__restrict__ chunks, int64_t n_chunks,
                                                    uint32_t *__restrict__ sums) {
    __shared__ uint32_t lds[4];
    const int64_t c = blockIdx.x;
    if (c >= n_chunks) return;
    const ChunkJob ch = chunks[c];
    const PlaneJob J = jobs[ch.job];
    if (!(J.flags & kDelta)) return;
    uint32_t acc = 0;
    const uint64_t e0 = ch.e0 + (uint64_t)threadIdx.x * 8u;
    for (int i = 0; i < 8; ++i) {
        const uint64_t e = e0 + i;
        if (e < J.n && e < ch.e0 + kPlaneChunk) acc = padd(acc, elem_word(stage + J.src, J.n, J.width, true, e));
    }
    uint32_t total;
    (void)block_scan_padd(acc, lds, total);
    if (threadIdx.x == 0) sums[c] = total;
}

// Exclusive packed scan of the chunk sums within each delta-coded job.
__global__ void k_plane_carry(const PlaneJob *__restrict__ jobs, int64_t n_jobs, const uint32_t *__restrict__ sums,
                              uint32_t *__restrict__ carry) {
    const int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= n_jobs) return;
    const PlaneJob J = jobs[j];
    if (!(J.flags & kDelta)) return;
    const uint64_t nc = (J.n + kPlaneChunk - 1) / kPlaneChunk;
    uint32_t run = 0;
    for (uint64_t c = 0; c < nc; ++c) {
        carry[J.chunk0 + c] = run;
        run = padd(run, sums[J.chunk0 + c]);
    }
}

// IEEE binary16 -> binary32, exact (subnormals scaled, Inf/NaN kept).
__device__ __forceinline__ float half_bits_to_float(uint32_t h) {
    const uint32_t sign = (h & 0x8000u) << 16, e = (h >> 10) & 31u, m = h & 1023u;
    if (e == 0) {
        const float v = (float)m * 5.9604644775390625e-8f;  // m * 2^-24
        return sign ? -v : v;
    }
    const uint32_t bits = e == 31 ? (sign | 0x7f800000u | (m << 13)) : (sign | ((e + 112u) << 23) | (m << 13));
    return __uint_as_float(bits);
}

// Un-delta (packed prefix sum + chunk carry), un-shuffle and widen.
__global__ __launch_bounds__(256) void k_plane_emit(const uint8_t *__restrict__ stage,
                                                    const PlaneJob *__restrict__ jobs,
                                                    const ChunkJob *__restrict__ chunks, int64_t n_chunks,
                                                    const uint32_t *__restrict__ carry, int32_t *__restrict__ idx_out,
                                                    float *__restrict__ val_out, uint32_t bound,
                                                    uint32_t *__restrict__ err) {
    __shared__ uint32_t lds[4];
    const int64_t c = blockIdx.x;
    if (c >= n_chunks) return;
    const ChunkJob ch = chunks[c];
    const PlaneJob J = jobs[ch.job];
    const bool pm = J.flags & kPlaneMajor, delta = J.flags & kDelta, is_idx = J.flags & kIndex;
    const uint32_t venc = (J.flags >> 8) & 0xffu;
    const uint64_t e0 = ch.e0 + (uint64_t)threadIdx.x * 8u;
    const uint64_t lim = J.n < ch.e0 + kPlaneChunk ? J.n : ch.e0 + kPlaneChunk;
    uint32_t w[8];
#pragma unroll
    for (int i = 0; i < 8; ++i) w[i] = (e0 + i < lim) ? elem_word(stage + J.src, J.n, J.width, pm, e0 + i) : 0u;
    if (delta) {
        uint32_t run = 0;
#pragma unroll
        for (int i = 0; i < 8; ++i) run = padd(run, w[i]);
        uint32_t total;
        const uint32_t incl = block_scan_padd(run, lds, total);
        // exclusive prefix of this thread = incl - own run, rebuilt by scanning
        uint32_t acc = padd(carry[c], incl);
        // acc now includes this thread's 8 words; walk them backwards
        uint32_t v[8];
#pragma unroll
        for (int i = 7; i >= 0; --i) {
            v[i] = acc;
            // acc -= w[i] (packed): add the packed two's complement
            const uint32_t neg = padd(~w[i], 0x01010101u);
            acc = padd(acc, neg);
        }
#pragma unroll
        for (int i = 0; i < 8; ++i) w[i] = v[i];
    }
    uint32_t bad = 0;
#pragma unroll
    for (int i = 0; i < 8; ++i) {
        const uint64_t e = e0 + i;
        if (e >= lim) break;
        const uint32_t x = w[i];
        if (is_idx) {
            if (x > 0x7fffffffu) bad |= ERR_I32;
            else if (bound && x >= bound) bad |= ERR_BOUND;
            idx_out[J.out + e] = (int32_t)x;
        } else {
            float f;
            if (venc == V_F32) f = __uint_as_float(x);
            else if (venc == V_F16) f = half_bits_to_float(x);
            else f = (float)x;
            val_out[J.out + e] = f;
        }
    }
    if (bad) atomicOr(err, bad);
}

// ---------------------------------------------------------------------------
// Host side
// ---------------------------------------------------------------------------
inline uint16_t rd16(const uint8_t *p) { return (uint16_t)(p[0] | (p[1] << 8)); }
inline uint32_t rd32(const uint8_t *p) {
    return (uint32_t)p[0] | ((uint32_t)p[1] << 8) | ((uint32_t)p[2] << 16) | ((uint32_t)p[3] << 24);
}
inline uint64_t rd64(const uint8_t *p) { return (uint64_t)rd32(p) | ((uint64_t)rd32(p + 4) << 32); }

[[noreturn]] inline void bad_shard(int64_t shard, const std::string &m) {
    throw Error(SCXPCA_ERR_INVALID, "shard " + std::to_string(shard) + ": " + m);
}

// The 76-byte header (scx-format/src/shard.rs:124-178).
struct Hdr {
    int version, type, codec, venc, u16;
    int64_t n_major, n_minor, nnz, global_offset;
    uint32_t off[8];  // indptr, indices, values, block index: (offset, length)
};

inline Hdr parse_header(const uint8_t *p, uint64_t len, int64_t shard) {
    if (!p) bad_shard(shard, "null shard pointer");
    if (len < (uint64_t)kHdr)
        bad_shard(shard, "shard too small: " + std::to_string(len) + " bytes < 76 header");
    if (std::memcmp(p, "SCXS", 4) != 0) bad_shard(shard, "shard header: bad magic");
    Hdr h;
    h.version = p[4];
    if (h.version > 2) bad_shard(shard, "unsupported shard format version " + std::to_string(h.version));
    h.type = p[5];
    h.codec = p[6];
    h.venc = p[7];
    h.u16 = p[8] == 0;
    h.n_major = rd32(p + 12);
    h.n_minor = rd32(p + 16);
    h.nnz = (int64_t)rd64(p + 20);
    h.global_offset = (int64_t)rd64(p + 28);
    for (int i = 0; i < 8; ++i) h.off[i] = rd32(p + 36 + 4 * i);
    static const char *names[4] = {"indptr", "indices", "values", "block_index"};
    for (int i = 0; i < 4; ++i)
        if ((uint64_t)h.off[2 * i] + h.off[2 * i + 1] > len)
            bad_shard(shard, std::string(names[i]) + " slice [" + std::to_string(h.off[2 * i]) + ".." +
                                 std::to_string((uint64_t)h.off[2 * i] + h.off[2 * i + 1]) + "] exceeds shard size " +
                                 std::to_string(len));
    if (h.codec > C_SHUF) bad_shard(shard, "unknown codec_id: " + std::to_string(h.codec));
    if (h.venc > V_F16) bad_shard(shard, "unknown value_encoding: " + std::to_string(h.venc));
    if (h.nnz < 0) bad_shard(shard, "nnz overflows");
    return h;
}

inline int venc_width(int v) { return v == V_U8 ? 1 : (v == V_U16 || v == V_F16) ? 2 : 4; }
inline bool venc_integer(int v) { return v <= V_U32; }

struct Group {
    int64_t shard;
    const Hdr *h;
    const uint8_t *ip, *ix, *vv;
    uint64_t nip, nix, nvv;
    int64_t n_rows, nnz;
    int64_t row0, nnz0;          // global row / nnz offsets
    uint64_t st_ix, st_vv;       // staging offsets
    uint64_t sz_ix, sz_vv;       // staged bytes
    std::vector<RowJob> rows;
    std::vector<RiceJob> rice, rice0;  // general blocks, k = 0 blocks
    std::vector<PlaneJob> planes;
    std::string err;
};

// Block index of a framed shard (resolve_block_index, scx-format/src/shard.rs:594-730).
inline void resolve_groups(const uint8_t *sh, const Hdr &h, int64_t shard, std::vector<Group> &out) {
    const uint8_t *ip = sh + h.off[0], *ix = sh + h.off[2], *vv = sh + h.off[4], *bi = sh + h.off[6];
    if (h.version < 2) {
        Group g{};
        g.shard = shard, g.h = &h, g.ip = ip, g.ix = ix, g.vv = vv;
        g.nip = h.off[1], g.nix = h.off[3], g.nvv = h.off[5];
        g.n_rows = h.n_major, g.nnz = h.nnz;
        out.push_back(std::move(g));
        return;
    }
    const uint64_t bl = h.off[7];
    if (bl < 4) bad_shard(shard, "framed block index: truncated");
    const uint64_t ng = rd32(bi);
    if (ng * kBie > bl - 4) bad_shard(shard, "framed block index: " + std::to_string(ng) + " entries exceed its length");
    if (ng == 0) {
        if (h.n_major == 0 && h.nnz == 0) return;
        bad_shard(shard, "empty block index, but the header declares n_major " + std::to_string(h.n_major) +
                             " / nnz " + std::to_string(h.nnz));
    }
    int64_t row = 0, nsum = 0;
    uint32_t prev[3] = {0, 0, 0};
    for (uint64_t i = 0; i < ng; ++i) {
        const uint8_t *e = bi + 4 + kBie * i;
        const int64_t rs = rd32(e), nr = rd16(e + 4), gnnz = rd32(e + 18);
        const uint32_t o[3] = {rd32(e + 6), rd32(e + 10), rd32(e + 14)};
        const std::string at = "framed block index entry " + std::to_string(i) + ": ";
        if (rs != row) bad_shard(shard, at + "row_start " + std::to_string(rs) + " != expected " + std::to_string(row));
        if (nr == 0) bad_shard(shard, at + "n_rows == 0");
        if (gnnz > nr * h.n_minor) bad_shard(shard, at + "nnz_in_block exceeds the group capacity");
        if (i == 0 && (o[0] | o[1] | o[2])) bad_shard(shard, at + "first entry sub-stream offsets must all be 0");
        if (o[0] < prev[0] || o[1] < prev[1] || o[2] < prev[2]) bad_shard(shard, at + "non-monotonic sub-stream offset");
        prev[0] = o[0], prev[1] = o[1], prev[2] = o[2];
        row += nr;
        nsum += gnnz;
    }
    if (row != h.n_major)
        bad_shard(shard, "framed block index covers " + std::to_string(row) + " rows, header n_major " +
                             std::to_string(h.n_major));
    if (nsum != h.nnz)
        bad_shard(shard, "framed block index nnz " + std::to_string(nsum) + " != header nnz " + std::to_string(h.nnz));
    for (uint64_t i = 0; i < ng; ++i) {
        const uint8_t *e = bi + 4 + kBie * i;
        uint32_t s[3] = {rd32(e + 6), rd32(e + 10), rd32(e + 14)}, t[3];
        if (i + 1 < ng) {
            const uint8_t *e2 = e + kBie;
            t[0] = rd32(e2 + 6), t[1] = rd32(e2 + 10), t[2] = rd32(e2 + 14);
        } else {
            t[0] = h.off[1], t[1] = h.off[3], t[2] = h.off[5];
        }
        const uint32_t lens[3] = {h.off[1], h.off[3], h.off[5]};
        for (int k = 0; k < 3; ++k)
            if (s[k] > t[k] || t[k] > lens[k])
                bad_shard(shard, "framed block index entry " + std::to_string(i) + ": sub-stream range invalid");
        const int64_t gnnz = rd32(e + 18);
        if (t[0] == s[0]) bad_shard(shard, "framed block index entry " + std::to_string(i) + ": empty indptr range");
        if (gnnz > 0 && (t[1] == s[1] || t[2] == s[2]))
            bad_shard(shard, "framed block index entry " + std::to_string(i) +
                                 ": empty indices/values range with nnz > 0");
        Group g{};
        g.shard = shard, g.h = &h;
        g.ip = ip + s[0], g.ix = ix + s[1], g.vv = vv + s[2];
        g.nip = t[0] - s[0], g.nix = t[1] - s[1], g.nvv = t[2] - s[2];
        g.n_rows = rd16(e + 4), g.nnz = gnnz;
        out.push_back(std::move(g));
    }
}

// ---- system frame libraries (loaded on first use) ----
struct FrameLibs {
    typedef size_t (*zd_t)(void *, size_t, const void *, size_t);
    typedef unsigned (*ze_t)(size_t);
    typedef size_t (*lmk_t)(void **, unsigned);
    typedef size_t (*lfr_t)(void *);
    typedef size_t (*ld_t)(void *, void *, size_t *, const void *, size_t *, const void *);
    typedef unsigned (*le_t)(size_t);
    zd_t zdec = nullptr;
    ze_t zerr = nullptr;
    lmk_t lmk = nullptr;
    lfr_t lfr = nullptr;
    ld_t ldec = nullptr;
    le_t lerr = nullptr;
    FrameLibs() {
        if (void *z = dlopen("libzstd.so.1", RTLD_NOW | RTLD_LOCAL)) {
            zdec = (zd_t)dlsym(z, "ZSTD_decompress");
            zerr = (ze_t)dlsym(z, "ZSTD_isError");
        }
        if (void *l = dlopen("liblz4.so.1", RTLD_NOW | RTLD_LOCAL)) {
            lmk = (lmk_t)dlsym(l, "LZ4F_createDecompressionContext");
            lfr = (lfr_t)dlsym(l, "LZ4F_freeDecompressionContext");
            ldec = (ld_t)dlsym(l, "LZ4F_decompress");
            lerr = (le_t)dlsym(l, "LZ4F_isError");
        }
    }
    static const FrameLibs &get() {
        static FrameLibs libs;
        return libs;
    }
};

// A zstd frame that must inflate to exactly `want` bytes
// (zstd_decode_bounded + expect_exact_len, scx-codec/src/codecs/zstd_codec.rs:63-80).
inline std::string zstd_exact(const uint8_t *in, uint64_t n, uint8_t *out, uint64_t want, const char *what) {
    const FrameLibs &L = FrameLibs::get();
    if (!L.zdec || !L.zerr) return "libzstd.so.1 is not available: zstd-coded shards cannot be decoded";
    uint8_t dummy;
    const size_t r = L.zdec(want ? out : &dummy, want ? want : 1, in, n);
    if (L.zerr(r)) return std::string(what) + ": zstd frame is corrupt or inflates past " + std::to_string(want) + " bytes";
    if (r != want)
        return std::string(what) + ": zstd frame inflates to " + std::to_string(r) + " bytes, expected " +
               std::to_string(want);
    return "";
}

// An LZ4 frame that must inflate to exactly `want` bytes
// (lz4_frame_decompress, scx-codec/src/codecs/lz4_shuffle.rs:55-72).
inline std::string lz4_exact(const uint8_t *in, uint64_t n, uint8_t *out, uint64_t want, const char *what) {
    const FrameLibs &L = FrameLibs::get();
    if (!L.lmk || !L.ldec || !L.lerr || !L.lfr)
        return "liblz4.so.1 is not available: lz4shuffle-coded shards cannot be decoded";
    void *ctx = nullptr;
    if (L.lerr(L.lmk(&ctx, 100))) return "LZ4 decompression context";
    uint64_t got = 0, used = 0;
    std::string e;
    uint8_t spill[64];
    for (;;) {
        size_t dst = got < want ? (size_t)(want - got) : sizeof spill;
        uint8_t *dp = got < want ? out + got : spill;
        size_t src = (size_t)(n - used);
        const size_t r = L.ldec(ctx, dp, &dst, in + used, &src, nullptr);
        if (L.lerr(r)) {
            e = std::string(what) + ": corrupt LZ4 frame";
            break;
        }
        if (dp == spill && dst > 0) {
            e = std::string(what) + ": LZ4 frame inflates past " + std::to_string(want) + " bytes";
            break;
        }
        got += dst;
        used += src;
        if (r == 0) break;
        if (src == 0 && dst == 0) {
            e = std::string(what) + ": truncated LZ4 frame";
            break;
        }
    }
    L.lfr(ctx);
    if (e.empty() && got != want)
        e = std::string(what) + ": LZ4 frame inflates to " + std::to_string(got) + " bytes, expected " +
            std::to_string(want);
    return e;
}

// ---- bit walks ----
// 64 bits of a byte stream starting at bit p; bytes past `len` read as zero.
inline uint64_t host_window(const uint8_t *s, uint64_t len, uint64_t p) {
    const uint64_t b = p >> 3;
    uint64_t w = 0;
    if (b + 8 <= len) {
        std::memcpy(&w, s + b, 8);
    } else if (b < len) {
        std::memcpy(&w, s + b, (size_t)(len - b));
    }
    return w >> (p & 7);
}

// Unary code at bit p: number of ones before the next zero; the position
// after the zero, or false when the stream ends first.
inline bool host_unary(const uint8_t *s, uint64_t len, uint64_t &p, uint64_t &q) {
    const uint64_t end = len * 8;
    q = 0;
    while (p < end) {
        const uint64_t w = host_window(s, len, p);
        const uint64_t valid = std::min<uint64_t>(64 - (p & 7), end - p);
        const uint64_t inv = ~w;
        const uint64_t t = inv ? (uint64_t)__builtin_ctzll(inv) : 64;
        if (t < valid) {
            q += t;
            p += t + 1;
            return true;
        }
        q += valid;
        p += valid;
    }
    return false;
}

// Position just past the m-th zero bit at or after bit p (k = 0 Rice blocks:
// each code ends with its terminating zero), or false if the stream ends.
inline bool host_skip_zeros(const uint8_t *s, uint64_t len, uint64_t &p, uint32_t m) {
    const uint64_t end = len * 8;
    while (m > 0 && p < end) {
        const uint64_t w = host_window(s, len, p);
        const uint64_t valid = std::min<uint64_t>(64 - (p & 7), end - p);
        uint64_t z = ~w;
        if (valid < 64) z &= (1ull << valid) - 1;
        const uint32_t c = (uint32_t)__builtin_popcountll(z);
        if (c < m) {
            m -= c;
            p += valid;
            continue;
        }
        // the m-th set bit of z: skip whole bytes by popcount, then bits
        uint32_t pos = 0;
        for (uint32_t c8; (c8 = (uint32_t)__builtin_popcountll(z & 0xffu)) < m; z >>= 8, pos += 8) m -= c8;
        for (;; z >>= 1, ++pos)
            if ((z & 1u) && --m == 0) break;
        p += (uint64_t)pos + 1;
        return true;
    }
    return m == 0;
}

// Delta-Golomb indptr (scx-codec/src/delta_golomb.rs:90-160).
inline std::string dg_decode(const uint8_t *s, uint64_t len, int64_t n, uint64_t *out) {
    if (n == 0) return "";
    if (len < 8) return "Delta-Golomb indptr shorter than 8 bytes";
    if ((uint64_t)n > len * 8) return "Delta-Golomb indptr: entries exceed the stream's capacity";
    out[0] = rd64(s);
    if (n == 1) return "";
    if (len < 9) return "Delta-Golomb indptr: missing k byte";
    const uint32_t k = s[8];
    if (k > 15) return "Delta-Golomb indptr: k > 15";
    const uint8_t *b = s + 9;
    const uint64_t bl = len - 9;
    uint64_t p = 0;
    for (int64_t i = 1; i < n; ++i) {
        uint64_t q;
        if (!host_unary(b, bl, p, q)) return "Delta-Golomb indptr: truncated unary code";
        uint64_t r = 0;
        if (k) {
            if (p + k > bl * 8) return "Delta-Golomb indptr: truncated remainder";
            r = host_window(b, bl, p) & ((1ull << k) - 1);
            p += k;
        }
        if (k && q > (~0ull >> k)) return "Delta-Golomb indptr: delta overflows u64";
        const uint64_t d = (q << k) | r;
        if (out[i - 1] > ~0ull - d) return "Delta-Golomb indptr: indptr overflows u64";
        out[i] = out[i - 1] + d;
    }
    return "";
}

// Bytes of `count` elements of `w` bytes, element-major -> plane-major or
// plane-major (after the wrapping un-delta) -> u64 values, for the indptr of
// the frame codecs (scx-codec/src/shuffle.rs:47-79, byte_delta.rs).
inline void planes_to_u64(uint8_t *planes, int64_t count, bool delta, uint64_t *out) {
    if (delta && count > 1)
        for (int p = 0; p < 8; ++p)
            for (int64_t e = 1; e < count; ++e)
                planes[p * count + e] = (uint8_t)(planes[p * count + e] + planes[p * count + e - 1]);
    for (int64_t e = 0; e < count; ++e) {
        uint64_t v = 0;
        for (int b = 0; b < 8; ++b) v |= (uint64_t)planes[(int64_t)b * count + e] << (8 * b);
        out[e] = v;
    }
}

// FOR-BP walk: one RowJob per non-empty row (scx-codec/src/forbp.rs:309-500).
inline std::string forbp_walk(Group &g, uint64_t stage_base) {
    const uint8_t *s = g.ix;
    const uint64_t len = g.nix;
    const bool u16 = g.h->u16;
    uint64_t pos = 0;
    int64_t rows_done = 0, out = 0;
    std::vector<uint32_t> rn;
    while (rows_done < g.n_rows) {
        if (pos + 6 > len) return "FOR-BP: truncated block header";
        const int64_t nb = rd16(s + pos + 4);
        pos += 6;
        if (nb > g.n_rows - rows_done) return "FOR-BP: a block declares more rows than remain";
        rn.resize((size_t)nb);
        for (int64_t r = 0; r < nb; ++r) {
            uint32_t v = 0;
            int shift = 0;
            for (;;) {
                if (pos >= len) return "FOR-BP: truncated row nnz varint";
                const uint8_t byte = s[pos++];
                if (shift == 28) {
                    if (byte > 0x0f) return "FOR-BP: row nnz varint overflows u32";
                    v |= (uint32_t)byte << 28;
                    break;
                }
                v |= (uint32_t)(byte & 0x7f) << shift;
                if (!(byte & 0x80)) break;
                shift += 7;
            }
            rn[(size_t)r] = v;
        }
        for (int64_t r = 0; r < nb; ++r) {
            const uint64_t nnz = rn[(size_t)r];
            if (nnz == 0) continue;
            uint32_t fmin;
            if (u16) {
                if (pos + 2 > len) return "FOR-BP: truncated frame_min";
                fmin = rd16(s + pos);
                pos += 2;
            } else {
                if (pos + 4 > len) return "FOR-BP: truncated frame_min";
                fmin = rd32(s + pos);
                pos += 4;
            }
            if (pos + 1 > len) return "FOR-BP: truncated frame_bits";
            const uint32_t fb = s[pos++];
            if (fb > 32) return "FOR-BP: frame_bits " + std::to_string(fb) + " > 32";
            if ((int64_t)(out + nnz) > g.nnz)
                return "FOR-BP: the rows decode to more than the declared " + std::to_string(g.nnz) + " indices";
            const bool bp4x = nnz >= (uint64_t)kSimd;
            uint64_t payload = 0;
            if (fb) {
                const uint64_t full = bp4x ? nnz / kSimd : 0;
                payload = full * fb * 16 + ((nnz - full * kSimd) * fb + 7) / 8;
            }
            if (pos + payload > len) return "FOR-BP: truncated row payload";
            RowJob j;
            j.src = stage_base + pos;
            j.out = (uint64_t)(g.nnz0 + out);
            j.nnz = (uint32_t)nnz;
            j.fmin = fmin;
            j.meta = fb | ((bp4x && fb) ? 256u : 0u);
            j.pad = 0;
            g.rows.push_back(j);
            pos += payload;
            out += (int64_t)nnz;
        }
        rows_done += nb;
    }
    if (out != g.nnz)
        return "FOR-BP: decoded " + std::to_string(out) + " indices but " + std::to_string(g.nnz) + " were declared";
    return "";
}

// Rice walk: one RiceJob per block (scx-codec/src/rice.rs:86-146).
inline std::string rice_walk(Group &g, uint64_t stage_base) {
    const uint8_t *s = g.vv;
    const uint64_t len = g.nvv, end = len * 8;
    uint64_t p = 0;
    int64_t done = 0;
    while (done < g.nnz) {
        const uint32_t m = (uint32_t)std::min<int64_t>(kBVal, g.nnz - done);
        if (p + 8 > end) return "Rice: truncated block header";
        const uint8_t hdr = s[p >> 3];
        if (hdr & 0xf0) {
            char buf[96];
            std::snprintf(buf, sizeof buf, "Rice block header has non-zero reserved high nibble: 0x%02x", hdr);
            return buf;
        }
        const uint32_t k = hdr & 0x0f;
        p += 8;
        RiceJob job{stage_base * 8 + p, (uint64_t)(g.nnz0 + done), m | (k << 16), 0};
        const uint64_t body0 = p;
        if (k == 0 && len < (1ull << 29)) {
            // a value is q + 1 with q < 2^32 - 1 whenever the stream holds fewer than 2^32 bits
            if (!host_skip_zeros(s, len, p, m)) return "Rice: truncated unary code";
        } else {
            const uint64_t mask = (1ull << k) - 1;
            for (uint32_t i = 0; i < m; ++i) {
                uint64_t q;
                if (!host_unary(s, len, p, q)) return "Rice: truncated unary code";
                uint64_t r = 0;
                if (k) {
                    if (p + k > end) return "Rice: truncated remainder";
                    r = host_window(s, len, p) & mask;
                    p += k;
                }
                if (q > (0xffffffffull >> k) || (((q << k) | r) + 1) > 0xffffffffull)
                    return "Rice value overflows u32 (corrupt stream)";
            }
        }
        job.nbits = (uint32_t)std::min<uint64_t>(p - body0, 0xffffffffull);
        (k == 0 && job.nbits < (1u << 30) ? g.rice0 : g.rice).push_back(job);
        p = (p + 7) & ~7ull;
        done += m;
    }
    return "";
}

// Pinned staging shared by every call (grown on demand; calls serialize).
struct Staging {
    std::mutex mu;
    void *host = nullptr;
    size_t bytes = 0;
    void ensure(size_t n) {
        if (n <= bytes && host) return;
        if (host) (void)hipHostFree(host);
        host = nullptr;
        bytes = 0;
        const size_t want = std::max<size_t>(n, 1 << 20);
        HIPC(hipHostMalloc(&host, want, hipHostMallocDefault));
        bytes = want;
    }
    static Staging &get() {
        static Staging s;
        return s;
    }
};

inline uint64_t align16(uint64_t x) { return (x + 15) & ~15ull; }

inline void parallel_for(int64_t n, const std::function<void(int64_t)> &fn) {
    const int64_t nt = std::min<int64_t>(n, std::max<int64_t>(1, std::min<unsigned>(16, std::thread::hardware_concurrency())));
    if (nt <= 1) {
        for (int64_t i = 0; i < n; ++i) fn(i);
        return;
    }
    std::atomic<int64_t> next{0};
    std::vector<std::thread> th;
    for (int64_t t = 0; t < nt; ++t)
        th.emplace_back([&] {
            for (int64_t i; (i = next.fetch_add(1)) < n;) fn(i);
        });
    for (auto &t : th) t.join();
}

}  // namespace shardgpu

// ---------------------------------------------------------------------------
// C ABI
// ---------------------------------------------------------------------------
extern "C" scxpca_status scxshard_read_info(const uint8_t *shard, uint64_t len, scxshard_info *out) {
    using namespace shardgpu;
    return guarded([&] {
        if (!out) throw Error(SCXPCA_ERR_INVALID, "null argument");
        const Hdr h = parse_header(shard, len, 0);
        std::vector<Group> groups;
        resolve_groups(shard, h, 0, groups);
        out->n_major = h.n_major;
        out->n_minor = h.n_minor;
        out->nnz = h.nnz;
        out->global_offset = h.global_offset;
        out->codec_id = h.codec;
        out->value_encoding = h.venc;
        out->index_dtype = h.u16 ? 0 : 1;
        out->shard_format_version = h.version;
        out->shard_type = h.type;
        out->n_groups = (int32_t)groups.size();
    });
}

extern "C" scxpca_status scxshard_decode_csr(const uint8_t *const *shards, const uint64_t *lens, int64_t n_shards,
                                             int64_t *indptr, int32_t *indices, float *data, int64_t n_rows,
                                             int64_t nnz, int32_t device, void *stream, scxshard_stats *stats) {
    using namespace shardgpu;
    using clk = std::chrono::steady_clock;
    return guarded([&] {
        const auto t0 = clk::now();
        if (!shards || !lens || !indptr || (nnz > 0 && (!indices || !data)))
            throw Error(SCXPCA_ERR_INVALID, "null argument");
        if (n_shards <= 0) throw Error(SCXPCA_ERR_INVALID, "decode_csr_shards_to_device: no CSR shards");
        DeviceGuard dg(device);
        // 1. headers, block indexes, groups
        std::vector<Hdr> hdr((size_t)n_shards);
        int64_t rows_sum = 0, nnz_sum = 0, n_cols = 0;
        for (int64_t s = 0; s < n_shards; ++s) {
            hdr[(size_t)s] = parse_header(shards[s], lens[s], s);
            const Hdr &h = hdr[(size_t)s];
            if (h.type != 0) bad_shard(s, "not a CSR shard (shard_type " + std::to_string(h.type) + ")");
            if (h.codec == C_PCODEC)
                bad_shard(s, "pcodec shards have no decoder in this library (codec_id 4)");
            if (h.codec == C_SCX1 && !venc_integer(h.venc))
                bad_shard(s, "Scx1 codec does not support float value encodings");
            if (s == 0) n_cols = h.n_minor;
            else if (h.n_minor != n_cols)
                bad_shard(s, "column count " + std::to_string(h.n_minor) + " != " + std::to_string(n_cols) +
                                 " (shards must agree)");
            rows_sum += h.n_major;
            nnz_sum += h.nnz;
        }
        if (rows_sum != n_rows || nnz_sum != nnz)
            throw Error(SCXPCA_ERR_INVALID, "output sized for " + std::to_string(n_rows) + " rows / " +
                                                std::to_string(nnz) + " nnz, the shards hold " +
                                                std::to_string(rows_sum) + " / " + std::to_string(nnz_sum));
        std::vector<Group> groups;
        for (int64_t s = 0; s < n_shards; ++s) resolve_groups(shards[s], hdr[(size_t)s], s, groups);
        // 2. global offsets and staging layout
        uint64_t stage = 0;
        {
            int64_t row = 0, nz = 0;
            for (auto &g : groups) {
                g.row0 = row, g.nnz0 = nz;
                row += g.n_rows, nz += g.nnz;
                const Hdr &h = *g.h;
                const uint64_t iw = h.u16 ? 2 : 4, vw = venc_width(h.venc);
                g.sz_ix = h.codec == C_SCX1 ? g.nix : (uint64_t)g.nnz * iw;
                g.sz_vv = h.codec == C_SCX1 ? g.nvv : (uint64_t)g.nnz * vw;
                g.st_ix = stage;
                stage = align16(stage + g.sz_ix + kStagePad);
            }
            // all value payloads after all index payloads: consecutive Rice
            // blocks stay adjacent across row groups and shards
            for (auto &g : groups) {
                g.st_vv = stage;
                stage = align16(stage + g.sz_vv + kStagePad);
            }
        }
        // 3. batches of groups (about a sixth of the payload each, at least
        //    8 MiB): the host work of batch b + 1 runs while batch b is copied
        //    and decoded on the stream.
        std::vector<size_t> cuts{0};
        {
            const uint64_t target = std::max<uint64_t>(stage / 6, 8ull << 20);
            uint64_t acc = 0;
            for (size_t i = 0; i < groups.size(); ++i) {
                acc += groups[i].sz_ix + groups[i].sz_vv;
                if (acc >= target && i + 1 < groups.size()) {
                    cuts.push_back(i + 1);
                    acc = 0;
                }
            }
            cuts.push_back(groups.size());
        }
        // work-list capacity per batch (upper bounds: a row job per row, a
        // Rice job per 256 values, two plane jobs and their chunks per group)
        struct Cap {
            uint64_t rows = 0, rice = 0, planes = 0, chunks = 0;
        };
        std::vector<Cap> cap(cuts.size() - 1);
        Cap tot;
        for (size_t b = 0; b + 1 < cuts.size(); ++b) {
            for (size_t i = cuts[b]; i < cuts[b + 1]; ++i) {
                const Group &g = groups[i];
                cap[b].rows += (uint64_t)g.n_rows;
                cap[b].rice += ((uint64_t)g.nnz + kBVal - 1) / kBVal;
                cap[b].planes += 2;
                cap[b].chunks += 2 * (((uint64_t)g.nnz + kPlaneChunk - 1) / kPlaneChunk);
            }
            tot.rows += cap[b].rows, tot.rice += cap[b].rice, tot.planes += cap[b].planes, tot.chunks += cap[b].chunks;
        }
        const uint64_t o_rows = align16(stage + 64), o_rice = align16(o_rows + tot.rows * sizeof(RowJob)),
                       o_pl = align16(o_rice + tot.rice * sizeof(RiceJob)),
                       o_ch = align16(o_pl + tot.planes * sizeof(PlaneJob)),
                       o_ip = align16(o_ch + tot.chunks * sizeof(ChunkJob)),
                       host_bytes = align16(o_ip + ((uint64_t)n_rows + 1) * 8),
                       o_sums = host_bytes, o_carry = align16(o_sums + tot.chunks * 4),
                       o_err = align16(o_carry + tot.chunks * 4), dev_bytes = o_err + 16;
        Staging &st = Staging::get();
        std::lock_guard<std::mutex> lock(st.mu);
        st.ensure(host_bytes);
        uint8_t *hs = static_cast<uint8_t *>(st.host);
        int64_t *h_indptr = reinterpret_cast<int64_t *>(hs + o_ip);
        h_indptr[0] = 0;
        hipStream_t sm = static_cast<hipStream_t>(stream);
        hipStream_t own = nullptr;
        if (!sm) {
            HIPC(hipStreamCreate(&own));
            sm = own;
        }
        struct StreamOwner {
            hipStream_t s;
            ~StreamOwner() {
                if (s) (void)hipStreamDestroy(s);
            }
        } so{own};
        DevBuf dbuf;
        dbuf.ensure(dev_bytes);
        uint8_t *d = dbuf.as<uint8_t>();
        const size_t nb = cuts.size() - 1;
        std::vector<hipEvent_t> ev(3 * nb, nullptr);  // per batch: copies start, copies end, kernels end
        struct EvOwner {
            std::vector<hipEvent_t> &e;
            ~EvOwner() {
                for (auto x : e)
                    if (x) (void)hipEventDestroy(x);
            }
        } eo{ev};
        for (auto &e : ev) HIPC(hipEventCreate(&e));
        uint32_t *err = reinterpret_cast<uint32_t *>(d + o_err);
        const uint32_t bound = (uint32_t)std::min<int64_t>(n_cols, 0xffffffffll);
        const uint32_t *st32 = reinterpret_cast<const uint32_t *>(d);
        uint64_t up_bytes = 0;
        double host_s = std::chrono::duration<double>(clk::now() - t0).count();
        bool launched = false;
        uint32_t herr = 0;
        try {
            Cap cur;  // table cursors
            for (size_t b = 0; b < nb; ++b) {
                const auto th0 = clk::now();
                // 3a. host work of the batch's groups
                parallel_for((int64_t)(cuts[b + 1] - cuts[b]), [&](int64_t li) {
                    Group &g = groups[cuts[b] + (size_t)li];
                    const Hdr &h = *g.h;
                    const int iw = h.u16 ? 2 : 4, vw = venc_width(h.venc);
                    try {
                        // indptr
                        std::vector<uint64_t> ip((size_t)g.n_rows + 1);
                        std::string e;
                        const uint64_t ipb = (uint64_t)(g.n_rows + 1) * 8;
                        if (h.codec == C_SCX1) {
                            e = dg_decode(g.ip, g.nip, g.n_rows + 1, ip.data());
                        } else if (h.codec == C_NONE) {
                            if (g.nip != ipb)
                                e = "indptr: " + std::to_string(g.nip) + " bytes, expected " + std::to_string(ipb);
                            else
                                for (int64_t r = 0; r <= g.n_rows; ++r) ip[(size_t)r] = rd64(g.ip + 8 * r);
                        } else {
                            std::vector<uint8_t> raw(ipb);
                            if (h.codec == C_ZSTD) {
                                e = zstd_exact(g.ip, g.nip, raw.data(), ipb, "indptr");
                                if (e.empty())
                                    for (int64_t r = 0; r <= g.n_rows; ++r) ip[(size_t)r] = rd64(raw.data() + 8 * r);
                            } else {
                                e = h.codec == C_LZ4 ? lz4_exact(g.ip, g.nip, raw.data(), ipb, "indptr")
                                                     : zstd_exact(g.ip, g.nip, raw.data(), ipb, "indptr");
                                if (e.empty()) planes_to_u64(raw.data(), g.n_rows + 1, h.codec == C_SHUF, ip.data());
                            }
                        }
                        if (e.empty()) {
                            // check_indptr_shape (scx-codec/src/guards.rs:258-305)
                            if (ip[0] != 0) e = "indptr must start at 0, got " + std::to_string(ip[0]);
                            for (int64_t r = 1; e.empty() && r <= g.n_rows; ++r)
                                if (ip[(size_t)r] < ip[(size_t)r - 1])
                                    e = "indptr not monotone at row " + std::to_string(r - 1);
                            if (e.empty() && ip[(size_t)g.n_rows] != (uint64_t)g.nnz)
                                e = "indptr ends at " + std::to_string(ip[(size_t)g.n_rows]) + ", nnz is " +
                                    std::to_string(g.nnz);
                        }
                        if (e.empty())
                            for (int64_t r = 1; r <= g.n_rows; ++r)
                                h_indptr[g.row0 + r] = g.nnz0 + (int64_t)ip[(size_t)r];
                        // payloads
                        if (e.empty() && g.nnz > 0) {
                            uint8_t *dix = hs + g.st_ix, *dvv = hs + g.st_vv;
                            if (h.codec == C_SCX1) {
                                if ((uint64_t)g.nnz > g.nix * 8 || (uint64_t)g.nnz > g.nvv * 8)
                                    e = "Scx1: declared nnz exceeds the sub-streams' capacity";
                                if (e.empty()) {
                                    std::memcpy(dix, g.ix, g.nix);
                                    std::memcpy(dvv, g.vv, g.nvv);
                                    e = forbp_walk(g, g.st_ix);
                                }
                                if (e.empty()) e = rice_walk(g, g.st_vv);
                            } else {
                                const bool shuffled = h.codec == C_LZ4 || h.codec == C_SHUF;
                                if (h.codec == C_NONE) {
                                    if (g.nix != g.sz_ix || g.nvv != g.sz_vv)
                                        e = "None codec: sub-stream lengths disagree with nnz";
                                    else {
                                        std::memcpy(dix, g.ix, g.nix);
                                        std::memcpy(dvv, g.vv, g.nvv);
                                    }
                                } else if (h.codec == C_LZ4) {
                                    e = lz4_exact(g.ix, g.nix, dix, g.sz_ix, "indices");
                                    if (e.empty()) e = lz4_exact(g.vv, g.nvv, dvv, g.sz_vv, "values");
                                } else {
                                    e = zstd_exact(g.ix, g.nix, dix, g.sz_ix, "indices");
                                    if (e.empty()) e = zstd_exact(g.vv, g.nvv, dvv, g.sz_vv, "values");
                                }
                                PlaneJob ji{g.st_ix, (uint64_t)g.nnz0, (uint64_t)g.nnz, (uint32_t)iw,
                                            kIndex | (shuffled && iw > 1 ? kPlaneMajor : 0u) |
                                                (h.codec == C_SHUF ? kDelta : 0u),
                                            0};
                                // float values of ShufDeltaZstd are zstd only (no shuffle)
                                const bool vshuf = h.codec == C_LZ4 || (h.codec == C_SHUF && venc_integer(h.venc));
                                PlaneJob jv{g.st_vv, (uint64_t)g.nnz0, (uint64_t)g.nnz, (uint32_t)vw,
                                            (vshuf && vw > 1 ? kPlaneMajor : 0u) | ((uint32_t)h.venc << 8), 0};
                                g.planes.push_back(ji);
                                g.planes.push_back(jv);
                            }
                        }
                        if (!e.empty()) g.err = "shard " + std::to_string(g.shard) + ": " + e;
                    } catch (const std::exception &ex) {
                        g.err = "shard " + std::to_string(g.shard) + ": " + ex.what();
                    }
                });
                for (size_t i = cuts[b]; i < cuts[b + 1]; ++i)
                    if (!groups[i].err.empty()) throw Error(SCXPCA_ERR_INVALID, groups[i].err);
                // 3b. the batch's work lists, in the pinned buffer (group
                //     offsets by a prefix sum, the copies in parallel)
                RowJob *rows = reinterpret_cast<RowJob *>(hs + o_rows) + cur.rows;
                RiceJob *rice = reinterpret_cast<RiceJob *>(hs + o_rice) + cur.rice;
                PlaneJob *planes = reinterpret_cast<PlaneJob *>(hs + o_pl) + cur.planes;
                ChunkJob *chunks = reinterpret_cast<ChunkJob *>(hs + o_ch) + cur.chunks;
                const size_t ng_b = cuts[b + 1] - cuts[b];
                std::vector<uint64_t> orow(ng_b + 1, 0), orc(ng_b + 1, 0), or0(ng_b + 1, 0), opl(ng_b + 1, 0),
                    och(ng_b + 1, 0);
                for (size_t li = 0; li < ng_b; ++li) {
                    const Group &g = groups[cuts[b] + li];
                    uint64_t nch = 0;
                    for (const PlaneJob &pj : g.planes) nch += (pj.n + kPlaneChunk - 1) / kPlaneChunk;
                    orow[li + 1] = orow[li] + g.rows.size();
                    orc[li + 1] = orc[li] + g.rice.size();
                    or0[li + 1] = or0[li] + g.rice0.size();
                    opl[li + 1] = opl[li] + g.planes.size();
                    och[li + 1] = och[li] + nch;
                }
                const uint64_t n_rows_b = orow[ng_b], n_rice = orc[ng_b], n_rice0 = or0[ng_b], n_pl = opl[ng_b],
                               n_ch = och[ng_b];
                parallel_for((int64_t)ng_b, [&](int64_t li) {
                    Group &g = groups[cuts[b] + (size_t)li];
                    if (!g.rows.empty()) std::memcpy(rows + orow[li], g.rows.data(), g.rows.size() * sizeof(RowJob));
                    if (!g.rice.empty()) std::memcpy(rice + orc[li], g.rice.data(), g.rice.size() * sizeof(RiceJob));
                    if (!g.rice0.empty())
                        std::memcpy(rice + n_rice + or0[li], g.rice0.data(), g.rice0.size() * sizeof(RiceJob));
                    uint64_t c = och[li], pl = opl[li];
                    for (PlaneJob pj : g.planes) {
                        pj.chunk0 = c;
                        for (uint64_t e0 = 0; e0 < pj.n; e0 += kPlaneChunk) chunks[c++] = ChunkJob{(uint32_t)pl, 0, e0};
                        planes[pl++] = pj;
                    }
                    std::vector<RowJob>().swap(g.rows);
                    std::vector<RiceJob>().swap(g.rice);
                    std::vector<RiceJob>().swap(g.rice0);
                });
                host_s += std::chrono::duration<double>(clk::now() - th0).count();
                // 3c. copies and kernels of the batch
                const Group &ga = groups[cuts[b]], &gz = groups[cuts[b + 1] - 1];
                const uint64_t ix0 = ga.st_ix, ix1 = gz.st_ix + gz.sz_ix + kStagePad;
                const uint64_t vv0 = ga.st_vv, vv1 = b + 1 == nb ? stage + 64 : gz.st_vv + gz.sz_vv + kStagePad;
                if (b + 1 == nb) std::memset(hs + stage, 0, 64);
                HIPC(hipEventRecord(ev[3 * b], sm));
                auto h2d = [&](uint64_t off, uint64_t bytes) {
                    if (!bytes) return;
                    HIPC(hipMemcpyAsync(d + off, hs + off, bytes, hipMemcpyHostToDevice, sm));
                    up_bytes += bytes;
                };
                h2d(ix0, ix1 - ix0);
                h2d(vv0, vv1 - vv0);
                h2d(o_rows + cur.rows * sizeof(RowJob), n_rows_b * sizeof(RowJob));
                h2d(o_rice + cur.rice * sizeof(RiceJob), (n_rice + n_rice0) * sizeof(RiceJob));
                h2d(o_pl + cur.planes * sizeof(PlaneJob), n_pl * sizeof(PlaneJob));
                h2d(o_ch + cur.chunks * sizeof(ChunkJob), n_ch * sizeof(ChunkJob));
                if (b == 0) {
                    k_zero_words<<<1, 64, 0, sm>>>(err, 4);
                    HIPC(hipGetLastError());
                }
                HIPC(hipEventRecord(ev[3 * b + 1], sm));
                launched = true;
                const auto *drows = reinterpret_cast<const RowJob *>(d + o_rows) + cur.rows;
                const auto *drice = reinterpret_cast<const RiceJob *>(d + o_rice) + cur.rice;
                if (n_rows_b)
                    k_forbp_rows<<<(unsigned)((n_rows_b + 3) / 4), 256, 0, sm>>>(st32, drows, (int64_t)n_rows_b,
                                                                              indices, bound, err);
                if (n_rice)
                    k_rice_blocks<<<(unsigned)((n_rice + 255) / 256), 256, 0, sm>>>(st32, drice, (int64_t)n_rice, data);
                if (n_rice0)
                    k_rice_k0<<<(unsigned)((n_rice0 + 3) / 4), 256, 0, sm>>>(st32, drice + n_rice, (int64_t)n_rice0,
                                                                          data);
                if (n_ch) {
                    const auto *pj = reinterpret_cast<const PlaneJob *>(d + o_pl) + cur.planes;
                    const auto *cj = reinterpret_cast<const ChunkJob *>(d + o_ch) + cur.chunks;
                    auto *sums = reinterpret_cast<uint32_t *>(d + o_sums) + cur.chunks;
                    auto *carry = reinterpret_cast<uint32_t *>(d + o_carry) + cur.chunks;
                    k_zero_words<<<blocks_for((int64_t)n_ch), 256, 0, sm>>>(carry, (int64_t)n_ch);
                    k_plane_sums<<<(unsigned)n_ch, 256, 0, sm>>>(d, pj, cj, (int64_t)n_ch, sums);
                    k_plane_carry<<<blocks_for((int64_t)n_pl), 256, 0, sm>>>(pj, (int64_t)n_pl, sums, carry);
                    k_plane_emit<<<(unsigned)n_ch, 256, 0, sm>>>(d, pj, cj, (int64_t)n_ch, carry, indices, data, bound,
                                                                 err);
                }
                HIPC(hipGetLastError());
                HIPC(hipEventRecord(ev[3 * b + 2], sm));
                cur.rows += cap[b].rows, cur.rice += cap[b].rice, cur.planes += cap[b].planes,
                    cur.chunks += cap[b].chunks;
            }
            HIPC(hipMemcpyAsync(indptr, h_indptr, ((size_t)n_rows + 1) * 8, hipMemcpyHostToDevice, sm));
            up_bytes += ((uint64_t)n_rows + 1) * 8;
            HIPC(hipMemcpyAsync(&herr, err, 4, hipMemcpyDeviceToHost, sm));
            HIPC(hipStreamSynchronize(sm));
        } catch (...) {
            // no DMA may still read the shared pinned buffer when the lock drops
            if (launched) (void)hipStreamSynchronize(sm);
            throw;
        }
        const auto t_end = clk::now();
        if (herr & ERR_U32) throw Error(SCXPCA_ERR_INVALID, "FOR-BP: a decoded index overflows u32 (corrupt stream)");
        if (herr & ERR_I32) throw Error(SCXPCA_ERR_INVALID, "a decoded index does not fit i32 (>= 2^31)");
        if (herr & ERR_BOUND)
            throw Error(SCXPCA_ERR_INVALID, "a decoded index is out of range for n_minor " + std::to_string(n_cols));
        if (stats) {
            double up_s = 0, dev_s = 0;
            for (size_t b = 0; b < nb; ++b) {
                float a = 0, c = 0;
                HIPC(hipEventElapsedTime(&a, ev[3 * b], ev[3 * b + 1]));
                HIPC(hipEventElapsedTime(&c, ev[3 * b + 1], ev[3 * b + 2]));
                up_s += a * 1e-3;
                dev_s += c * 1e-3;
            }
            std::memset(stats, 0, sizeof *stats);
            stats->host_uploaded_bytes = up_bytes;
            stats->device_decoded_bytes = (uint64_t)nnz * 8;
            stats->fully_device_decoded = 1;
            for (int64_t s = 0; s < n_shards; ++s) {
                const int c = hdr[(size_t)s].codec;
                if (c == C_SCX1) stats->n_shards_scx1_gpu++;
                else if (c == C_SHUF) stats->n_shards_shufdelta_gpu++;
                else stats->n_shards_other_gpu++;
                if (c != C_SCX1 && c != C_NONE) stats->fully_device_decoded = 0;
            }
            stats->n_groups = (uint32_t)groups.size();
            stats->host_seconds = host_s;
            stats->device_seconds = dev_s;
            stats->upload_seconds = up_s;
            stats->total_seconds = std::chrono::duration<double>(t_end - t0).count();
        }
    });
}
