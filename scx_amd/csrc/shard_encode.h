// Host writer of SCX CSR shard sections (C ABI scxshard_encode in
// include/scxshard.h): the counterpart of the device decoder, so a CSR can be
// written in the layout the reference writer produces and read back on the
// GPU. Codecs Scx1 (Delta-Golomb indptr, FOR-BP indices with the BitPacker4x
// layout for rows of >= 128 entries, adaptive Rice values) and None.
// Follows scx-codec/src/{rice.rs:34-80, delta_golomb.rs:31-82,
// forbp.rs:86-200}, scx-format/src/shard.rs:98-122 (header) and :410-478
// (block index), scx-format-io/src/encoder.rs:524-690 (row-group framing).
// Included by scxpca.hip after shard_decode.h.
#pragma once

#include <cmath>

namespace shardenc {

// LSB-first bit writer over a growable byte vector (64-bit accumulator).
struct Bits {
    std::vector<uint8_t> &out;
    uint64_t acc = 0;
    int n = 0;
    explicit Bits(std::vector<uint8_t> &o) : out(o) {}
    void put(uint64_t v, int bits) {  // bits <= 32
        if (!bits) return;
        acc |= (v & ((bits == 64) ? ~0ull : ((1ull << bits) - 1))) << n;
        n += bits;
        while (n >= 8) {
            out.push_back((uint8_t)acc);
            acc >>= 8;
            n -= 8;
        }
    }
    void unary(uint64_t q) {
        while (q >= 32) {
            put(0xffffffffu, 32);
            q -= 32;
        }
        put((1ull << q) - 1, (int)q);
        put(0, 1);
    }
    void pad() {
        if (n) {
            out.push_back((uint8_t)acc);
            acc = 0;
            n = 0;
        }
    }
};

inline int rice_k(uint64_t median) {
    if (!median) return 0;
    const double raw = std::floor(std::log2(0.69314718055994530942 * (double)median));
    return raw < 0 ? 0 : raw > 15 ? 15 : (int)raw;
}

template <typename T>
inline T floor_median(std::vector<T> &v) {
    if (v.empty()) return 0;
    if (v.size() == 1) return v[0];
    std::nth_element(v.begin(), v.begin() + (v.size() - 1) / 2, v.end());
    return v[(v.size() - 1) / 2];
}

inline void rice(const uint32_t *v, int64_t n, std::vector<uint8_t> &out) {
    Bits b(out);
    std::vector<uint32_t> tmp;
    for (int64_t b0 = 0; b0 < n; b0 += shardgpu::kBVal) {
        const int64_t m = std::min<int64_t>(shardgpu::kBVal, n - b0);
        tmp.assign(v + b0, v + b0 + m);
        for (auto &t : tmp) {
            if (t == 0) throw Error(SCXPCA_ERR_INVALID, "Scx1 values must be non-zero counts");
            t -= 1;
        }
        const int k = rice_k(floor_median(tmp));
        b.put((uint64_t)k, 8);
        for (int64_t i = 0; i < m; ++i) {
            const uint64_t s = v[b0 + i] - 1u;
            b.unary(s >> k);
            b.put(s, k);
        }
        b.pad();
    }
}

inline void delta_golomb(const uint64_t *ip, int64_t n, std::vector<uint8_t> &out) {
    if (n == 0) return;
    for (int i = 0; i < 8; ++i) out.push_back((uint8_t)(ip[0] >> (8 * i)));
    if (n == 1) return;
    std::vector<uint64_t> d((size_t)n - 1);
    for (int64_t i = 0; i + 1 < n; ++i) {
        if (ip[i + 1] < ip[i]) throw Error(SCXPCA_ERR_INVALID, "indptr must be monotone non-decreasing");
        d[(size_t)i] = ip[i + 1] - ip[i];
    }
    const int k = rice_k(floor_median(d));
    out.push_back((uint8_t)k);
    Bits b(out);
    for (int64_t i = 0; i + 1 < n; ++i) {
        const uint64_t dd = ip[i + 1] - ip[i];
        b.unary(dd >> k);
        b.put(dd, k);
    }
    b.pad();
}

// 128 deltas in the 4-lane layout: value p in lane p & 3, slot p >> 2,
// lane l's stream the u32 words at byte 16 w + 4 l.
inline void bp4x(const uint32_t *d, int fb, uint8_t *dst) {
    std::memset(dst, 0, (size_t)fb * 16);
    for (int l = 0; l < 4; ++l) {
        uint64_t acc = 0;
        int nb = 0, w = 0;
        for (int j = 0; j < 32; ++j) {
            acc |= (uint64_t)(d[4 * j + l] & (fb >= 32 ? 0xffffffffu : ((1u << fb) - 1))) << nb;
            nb += fb;
            while (nb >= 32) {
                std::memcpy(dst + 16 * w + 4 * l, &acc, 4);
                acc >>= 32;
                nb -= 32;
                ++w;
            }
        }
    }
}

inline void varint(uint32_t v, std::vector<uint8_t> &out) {
    do {
        uint8_t byte = v & 0x7f;
        v >>= 7;
        if (v) byte |= 0x80;
        out.push_back(byte);
    } while (v);
}

inline void forbp(const uint32_t *idx, const uint64_t *ip, int64_t n_rows, bool u16, std::vector<uint8_t> &out) {
    std::vector<uint32_t> d;
    for (int64_t b0 = 0; b0 < n_rows; b0 += shardgpu::kBIdx) {
        const int64_t nb = std::min<int64_t>(shardgpu::kBIdx, n_rows - b0);
        const uint64_t bnnz = ip[b0 + nb] - ip[b0];
        if (bnnz > 0xffffffffull) throw Error(SCXPCA_ERR_INVALID, "FOR-BP block nnz overflows u32");
        for (int i = 0; i < 4; ++i) out.push_back((uint8_t)(bnnz >> (8 * i)));
        out.push_back((uint8_t)nb);
        out.push_back((uint8_t)(nb >> 8));
        for (int64_t r = 0; r < nb; ++r) varint((uint32_t)(ip[b0 + r + 1] - ip[b0 + r]), out);
        for (int64_t r = 0; r < nb; ++r) {
            const uint64_t a = ip[b0 + r], nnz = ip[b0 + r + 1] - a;
            if (!nnz) continue;
            const uint32_t *ri = idx + a;
            d.assign((size_t)nnz, 0u);
            uint32_t mx = 0;
            for (uint64_t j = 1; j < nnz; ++j) {
                if (ri[j] < ri[j - 1]) throw Error(SCXPCA_ERR_INVALID, "FOR-BP requires sorted row indices");
                d[(size_t)j] = ri[j] - ri[j - 1];
                mx = std::max(mx, d[(size_t)j]);
            }
            const int fb = mx ? 32 - __builtin_clz(mx) : 0;
            if (u16) {
                if (ri[0] > 0xffff) throw Error(SCXPCA_ERR_INVALID, "frame_min above u16 with u16 indices");
                out.push_back((uint8_t)ri[0]);
                out.push_back((uint8_t)(ri[0] >> 8));
            } else {
                for (int i = 0; i < 4; ++i) out.push_back((uint8_t)(ri[0] >> (8 * i)));
            }
            out.push_back((uint8_t)fb);
            if (!fb) continue;
            uint64_t head = 0;
            if (nnz >= (uint64_t)shardgpu::kSimd) {
                const uint64_t full = nnz / shardgpu::kSimd;
                for (uint64_t c = 0; c < full; ++c) {
                    const size_t at = out.size();
                    out.resize(at + (size_t)fb * 16);
                    bp4x(d.data() + c * shardgpu::kSimd, fb, out.data() + at);
                }
                head = full * shardgpu::kSimd;
            }
            if (head < nnz) {
                Bits b(out);
                for (uint64_t j = head; j < nnz; ++j) b.put(d[(size_t)j], fb);
                b.pad();
            }
        }
    }
}

struct Streams {
    std::vector<uint8_t> ip, ix, vv;
};

inline void encode_group(const uint64_t *ip, int64_t n_rows, const uint32_t *idx, const uint8_t *vals, int codec,
                         int venc, bool u16, Streams &s) {
    const int64_t nnz = (int64_t)(ip[n_rows] - ip[0]);
    const int vw = shardgpu::venc_width(venc);
    if (codec == shardgpu::C_SCX1) {
        delta_golomb(ip, n_rows + 1, s.ip);
        forbp(idx, ip, n_rows, u16, s.ix);
        std::vector<uint32_t> v((size_t)nnz);
        for (int64_t i = 0; i < nnz; ++i)
            v[(size_t)i] = vw == 1 ? vals[i] : vw == 2 ? shardgpu::rd16(vals + 2 * i) : shardgpu::rd32(vals + 4 * i);
        rice(v.data(), nnz, s.vv);
        return;
    }
    for (int64_t i = 0; i <= n_rows; ++i)
        for (int b = 0; b < 8; ++b) s.ip.push_back((uint8_t)(ip[i] >> (8 * b)));
    for (int64_t i = 0; i < nnz; ++i)
        for (int b = 0; b < (u16 ? 2 : 4); ++b) s.ix.push_back((uint8_t)(idx[i] >> (8 * b)));
    s.vv.assign(vals, vals + (size_t)nnz * vw);
}

inline void put32(std::vector<uint8_t> &o, size_t at, uint32_t v) {
    for (int i = 0; i < 4; ++i) o[at + i] = (uint8_t)(v >> (8 * i));
}

}  // namespace shardenc

extern "C" scxpca_status scxshard_encode(const int64_t *indptr, int64_t n_rows, const uint32_t *indices,
                                         const void *values, int32_t value_encoding, int32_t codec,
                                         int32_t index_dtype, int64_t n_minor, int64_t global_offset,
                                         int32_t row_group_rows, uint8_t *out, uint64_t cap, uint64_t *written) {
    using namespace shardenc;
    return guarded([&] {
        if (!indptr || !written || (n_rows > 0 && indptr[n_rows] > indptr[0] && (!indices || !values)))
            throw Error(SCXPCA_ERR_INVALID, "null argument");
        if (codec != shardgpu::C_SCX1 && codec != shardgpu::C_NONE)
            throw Error(SCXPCA_ERR_INVALID, "scxshard_encode writes codec scx1 (1) or none (0)");
        if (value_encoding < 0 || value_encoding > shardgpu::V_F16)
            throw Error(SCXPCA_ERR_INVALID, "unknown value_encoding: " + std::to_string(value_encoding));
        if (codec == shardgpu::C_SCX1 && !shardgpu::venc_integer(value_encoding))
            throw Error(SCXPCA_ERR_INVALID, "Scx1 codec does not support float value encodings");
        if (n_rows < 0 || n_rows > 0xffffffffll || n_minor < 0 || n_minor > 0xffffffffll)
            throw Error(SCXPCA_ERR_INVALID, "shard dimensions exceed u32");
        const bool framed = row_group_rows > 0, u16 = index_dtype == 0;
        if (framed && n_rows == 0) throw Error(SCXPCA_ERR_INVALID, "a framed shard needs at least one row");
        const int vw = shardgpu::venc_width(value_encoding);
        std::vector<uint64_t> ip((size_t)n_rows + 1);
        for (int64_t r = 0; r <= n_rows; ++r) {
            if (indptr[r] < indptr[0] || (r && indptr[r] < indptr[r - 1]))
                throw Error(SCXPCA_ERR_INVALID, "indptr must be monotone non-decreasing");
            ip[(size_t)r] = (uint64_t)(indptr[r] - indptr[0]);
        }
        const uint32_t *idx = indices ? indices + indptr[0] : nullptr;
        const uint8_t *vals = values ? static_cast<const uint8_t *>(values) + (size_t)indptr[0] * vw : nullptr;
        const int64_t g = framed ? std::min<int64_t>(row_group_rows, 65535) : std::max<int64_t>(n_rows, 1);
        const int64_t ng = framed ? (n_rows + g - 1) / g : 1;
        std::vector<Streams> gs((size_t)ng);
        shardgpu::parallel_for(ng, [&](int64_t gi) {
            const int64_t r0 = framed ? gi * g : 0, r1 = framed ? std::min(n_rows, r0 + g) : n_rows;
            std::vector<uint64_t> loc((size_t)(r1 - r0 + 1));
            for (int64_t r = r0; r <= r1; ++r) loc[(size_t)(r - r0)] = ip[(size_t)r] - ip[(size_t)r0];
            encode_group(loc.data(), r1 - r0, idx + ip[(size_t)r0], vals + ip[(size_t)r0] * vw, codec,
                         value_encoding, u16, gs[(size_t)gi]);
        });
        const int64_t nent = framed ? ng : (n_rows > 65535 ? (n_rows + 65534) / 65535 : 1);
        uint64_t lip = 0, lix = 0, lvv = 0;
        for (auto &s : gs) lip += s.ip.size(), lix += s.ix.size(), lvv += s.vv.size();
        const uint64_t bil = 4 + (uint64_t)shardgpu::kBie * nent;
        const uint64_t total = shardgpu::kHdr + lip + lix + lvv + bil;
        if (lip + lix + lvv + bil > 0xffffffffull) throw Error(SCXPCA_ERR_INVALID, "shard sub-streams exceed u32 offsets");
        *written = total;
        if (!out) return;
        if (cap < total) throw Error(SCXPCA_ERR_INVALID, "output buffer too small: " + std::to_string(total) + " bytes needed");
        std::vector<uint8_t> o((size_t)total, 0);
        std::memcpy(o.data(), "SCXS", 4);
        o[4] = (uint8_t)(framed ? 2 : 1);
        o[6] = (uint8_t)codec;
        o[7] = (uint8_t)value_encoding;
        o[8] = (uint8_t)(u16 ? 0 : 1);
        put32(o, 12, (uint32_t)n_rows);
        put32(o, 16, (uint32_t)n_minor);
        const uint64_t nnz = ip[(size_t)n_rows];
        put32(o, 20, (uint32_t)nnz);
        put32(o, 24, (uint32_t)(nnz >> 32));
        put32(o, 28, (uint32_t)global_offset);
        put32(o, 32, (uint32_t)((uint64_t)global_offset >> 32));
        const uint32_t o_ip = shardgpu::kHdr, o_ix = o_ip + (uint32_t)lip, o_vv = o_ix + (uint32_t)lix,
                       o_bi = o_vv + (uint32_t)lvv;
        put32(o, 36, o_ip), put32(o, 40, (uint32_t)lip), put32(o, 44, o_ix), put32(o, 48, (uint32_t)lix);
        put32(o, 52, o_vv), put32(o, 56, (uint32_t)lvv), put32(o, 60, o_bi), put32(o, 64, (uint32_t)bil);
        // bytes 68..76: the payload checksum (BLAKE3, truncated) -- left zero; readers do not verify it
        uint64_t a = 0, b = 0, c = 0;
        std::vector<uint64_t> offs((size_t)ng * 3);
        for (int64_t gi = 0; gi < ng; ++gi) {
            const Streams &s = gs[(size_t)gi];
            offs[(size_t)gi * 3] = a, offs[(size_t)gi * 3 + 1] = b, offs[(size_t)gi * 3 + 2] = c;
            if (!s.ip.empty()) std::memcpy(o.data() + o_ip + a, s.ip.data(), s.ip.size());
            if (!s.ix.empty()) std::memcpy(o.data() + o_ix + b, s.ix.data(), s.ix.size());
            if (!s.vv.empty()) std::memcpy(o.data() + o_vv + c, s.vv.data(), s.vv.size());
            a += s.ip.size(), b += s.ix.size(), c += s.vv.size();
        }
        put32(o, o_bi, (uint32_t)nent);
        const int64_t step = framed ? g : 65535;
        for (int64_t e = 0; e < nent; ++e) {
            const size_t at = o_bi + 4 + (size_t)shardgpu::kBie * e;
            const int64_t r0 = e * step, r1 = std::min(n_rows, r0 + step);
            put32(o, at, (uint32_t)r0);
            o[at + 4] = (uint8_t)(r1 - r0);
            o[at + 5] = (uint8_t)((r1 - r0) >> 8);
            put32(o, at + 6, framed ? (uint32_t)offs[(size_t)e * 3] : 0u);
            put32(o, at + 10, framed ? (uint32_t)offs[(size_t)e * 3 + 1] : 0u);
            put32(o, at + 14, framed ? (uint32_t)offs[(size_t)e * 3 + 2] : 0u);
            put32(o, at + 18, (uint32_t)(ip[(size_t)r1] - ip[(size_t)r0]));
        }
        std::memcpy(out, o.data(), (size_t)total);
    });
}
