/* CPU oracle for the SCX shard codecs -- test infrastructure only.
 *
 * Each function restates the scalar reference in scx-codec / scx-format and
 * cites the lines it follows. The oracle is the checker for the GPU shard
 * decoder (scx_amd/csrc/shard_decode.hip); the product never links it. */
#define _GNU_SOURCE
#include "scxshard_oracle.h"

#include <dlfcn.h>
#include <math.h>
#include <stdarg.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

static __thread char g_err[512];

static int fail(const char *fmt, ...) {
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_err, sizeof g_err, fmt, ap);
    va_end(ap);
    return -1;
}

const char *sho_last_error(void) { return g_err; }

/* ------------------------------------------------------------------------
 * LSB-first bit writer / reader (scx-codec/src/bitstream.rs:12-160 writer,
 * :160-350 reader): bit i of the stream is bit (i % 8) of byte i / 8.
 * ------------------------------------------------------------------------ */
typedef struct {
    uint8_t *buf;
    int64_t cap, nbit;
    int overflow;
} bw_t;

/* a writer with buf == NULL only counts bits */
static void bw_bit(bw_t *w, int bit) {
    int64_t byte = w->nbit >> 3;
    if (!w->buf) {
        w->nbit++;
        return;
    }
    if (byte >= w->cap) {
        w->overflow = 1;
        w->nbit++;
        return;
    }
    if ((w->nbit & 7) == 0) w->buf[byte] = 0;
    if (bit) w->buf[byte] |= (uint8_t)(1u << (w->nbit & 7));
    w->nbit++;
}

static void bw_bits(bw_t *w, uint64_t v, int n) {
    for (int i = 0; i < n; ++i) bw_bit(w, (int)((v >> i) & 1u));
}

/* write_unary: q ones then a zero (bitstream.rs:112-124) */
static void bw_unary(bw_t *w, uint64_t q) {
    for (uint64_t i = 0; i < q; ++i) bw_bit(w, 1);
    bw_bit(w, 0);
}

static void bw_pad(bw_t *w) {
    while (w->nbit & 7) bw_bit(w, 0);
}

static int64_t bw_bytes(const bw_t *w) { return (w->nbit + 7) >> 3; }

typedef struct {
    const uint8_t *buf;
    int64_t len, pos; /* pos in bits */
} br_t;

static int br_bit(br_t *r, int *bit) {
    if ((r->pos >> 3) >= r->len) return -1;
    *bit = (r->buf[r->pos >> 3] >> (r->pos & 7)) & 1;
    r->pos++;
    return 0;
}

/* read_bits (bitstream.rs:246-285): n <= 64, LSB first */
static int br_bits(br_t *r, int n, uint64_t *v) {
    uint64_t x = 0;
    for (int i = 0; i < n; ++i) {
        int b;
        if (br_bit(r, &b)) return -1;
        x |= (uint64_t)b << i;
    }
    *v = x;
    return 0;
}

/* read_unary (bitstream.rs:288-320): count ones until a zero */
static int br_unary(br_t *r, uint64_t *q) {
    uint64_t c = 0;
    for (;;) {
        int b;
        if (br_bit(r, &b)) return -1;
        if (!b) break;
        ++c;
    }
    *q = c;
    return 0;
}

/* align_to_byte (bitstream.rs:338-350) */
static void br_align(br_t *r) { r->pos = (r->pos + 7) & ~(int64_t)7; }

/* ------------------------------------------------------------------------
 * floor median: the ((n - 1) / 2)-th smallest (scx-codec/src/median.rs:
 * floor_median_inplace_impl); k = clamp(floor(log2(ln2 * median)), 0, 15)
 * (rice.rs:27-33, delta_golomb.rs:17-23).
 * ------------------------------------------------------------------------ */
static int cmp_u64(const void *a, const void *b) {
    uint64_t x = *(const uint64_t *)a, y = *(const uint64_t *)b;
    return x < y ? -1 : x > y;
}

static uint64_t floor_median(uint64_t *v, int64_t n) {
    if (n == 0) return 0;
    if (n == 1) return v[0];
    qsort(v, (size_t)n, sizeof(uint64_t), cmp_u64);
    return v[(n - 1) / 2];
}

static int compute_k(uint64_t median) {
    if (median == 0) return 0;
    double raw = floor(log2(M_LN2 * (double)median));
    if (raw < 0) return 0;
    if (raw > 15) return 15;
    return (int)raw;
}

/* ------------------------------------------------------------------------
 * Adaptive Rice values (scx-codec/src/rice.rs:34-80 encode, :86-146 decode):
 * blocks of 256; a header byte with k in the low nibble (high nibble zero);
 * per value s = v - 1: unary(s >> k) then the k low bits; each block padded
 * to a byte.
 * ------------------------------------------------------------------------ */
#define B_VAL 256

int64_t sho_rice_encode(const uint32_t *v, int64_t n, uint8_t *out, int64_t cap) {
    bw_t w = {out, cap, 0, 0};
    uint64_t tmp[B_VAL];
    for (int64_t b0 = 0; b0 < n; b0 += B_VAL) {
        int64_t m = n - b0 < B_VAL ? n - b0 : B_VAL;
        for (int64_t i = 0; i < m; ++i) {
            if (v[b0 + i] == 0) return fail("rice_encode: zero value");
            tmp[i] = v[b0 + i] - 1u;
        }
        int k = compute_k(floor_median(tmp, m));
        bw_bits(&w, (uint64_t)k, 8);
        for (int64_t i = 0; i < m; ++i) {
            uint64_t s = v[b0 + i] - 1u;
            bw_unary(&w, s >> k);
            if (k > 0) bw_bits(&w, s & ((1ull << k) - 1), k);
        }
        bw_pad(&w);
    }
    if (w.overflow) return fail("rice_encode: output capacity");
    return bw_bytes(&w);
}

int64_t sho_rice_decode(const uint8_t *in, int64_t len, int64_t n, uint32_t *out) {
    br_t r = {in, len, 0};
    int64_t done = 0;
    while (done < n) {
        int64_t m = n - done < B_VAL ? n - done : B_VAL;
        uint64_t hdr;
        if (br_bits(&r, 8, &hdr)) return fail("rice: truncated block header");
        if (hdr & 0xF0) return fail("rice: reserved high nibble 0x%02x", (unsigned)hdr);
        int k = (int)(hdr & 0x0F);
        for (int64_t i = 0; i < m; ++i) {
            uint64_t q, rem = 0;
            if (br_unary(&r, &q)) return fail("rice: truncated unary code");
            if (k > 0 && br_bits(&r, k, &rem)) return fail("rice: truncated remainder");
            if (q > (0xFFFFFFFFull >> k)) return fail("rice: value overflows u32");
            uint64_t val = ((q << k) | rem) + 1;
            if (val > 0xFFFFFFFFull) return fail("rice: value overflows u32");
            out[done + i] = (uint32_t)val;
        }
        br_align(&r);
        done += m;
    }
    return (r.pos + 7) >> 3;
}

/* ------------------------------------------------------------------------
 * Delta-Golomb-Rice indptr (scx-codec/src/delta_golomb.rs:31-82 encode,
 * :90-160 decode): raw LE u64 first value, one k byte, Rice-coded deltas in a
 * single stream, padded at the end.
 * ------------------------------------------------------------------------ */
static void put_u64(uint8_t *p, uint64_t v) {
    for (int i = 0; i < 8; ++i) p[i] = (uint8_t)(v >> (8 * i));
}
static uint64_t get_u64(const uint8_t *p) {
    uint64_t v = 0;
    for (int i = 0; i < 8; ++i) v |= (uint64_t)p[i] << (8 * i);
    return v;
}
static uint32_t get_u32(const uint8_t *p) {
    return (uint32_t)p[0] | ((uint32_t)p[1] << 8) | ((uint32_t)p[2] << 16) | ((uint32_t)p[3] << 24);
}
static uint16_t get_u16(const uint8_t *p) { return (uint16_t)(p[0] | (p[1] << 8)); }
static void put_u32(uint8_t *p, uint32_t v) {
    for (int i = 0; i < 4; ++i) p[i] = (uint8_t)(v >> (8 * i));
}
static void put_u16(uint8_t *p, uint16_t v) {
    p[0] = (uint8_t)v;
    p[1] = (uint8_t)(v >> 8);
}

int64_t sho_dg_encode(const uint64_t *indptr, int64_t n, uint8_t *out, int64_t cap) {
    if (n == 0) return 0;
    if (out && cap < 8) return fail("dg_encode: output capacity");
    if (out) put_u64(out, indptr[0]);
    if (n == 1) return 8;
    uint64_t *d = malloc(sizeof(uint64_t) * (size_t)(n - 1));
    for (int64_t i = 0; i + 1 < n; ++i) {
        if (indptr[i + 1] < indptr[i]) {
            free(d);
            return fail("indptr must be monotone non-decreasing");
        }
        d[i] = indptr[i + 1] - indptr[i];
    }
    int k = compute_k(floor_median(d, n - 1));
    free(d);
    if (out && cap < 9) return fail("dg_encode: output capacity");
    if (out) out[8] = (uint8_t)k;
    bw_t w = {out ? out + 9 : NULL, cap - 9, 0, 0};
    for (int64_t i = 0; i + 1 < n; ++i) {
        uint64_t dd = indptr[i + 1] - indptr[i];
        bw_unary(&w, dd >> k);
        if (k > 0) bw_bits(&w, dd & ((1ull << k) - 1), k);
    }
    if (w.overflow) return fail("dg_encode: output capacity");
    return 9 + bw_bytes(&w);
}

int64_t sho_dg_decode(const uint8_t *in, int64_t len, int64_t n, uint64_t *out) {
    if (n == 0) return 0;
    if (len < 8) return fail("delta-golomb: stream shorter than 8 bytes");
    if (n > len * 8) return fail("delta-golomb: %lld entries exceed the stream's capacity", (long long)n);
    out[0] = get_u64(in);
    if (n == 1) return 8;
    if (len < 9) return fail("delta-golomb: missing k byte");
    int k = in[8];
    if (k > 15) return fail("delta-golomb: k %d > 15", k);
    br_t r = {in + 9, len - 9, 0};
    for (int64_t i = 1; i < n; ++i) {
        uint64_t q, rem = 0;
        if (br_unary(&r, &q)) return fail("delta-golomb: truncated unary code");
        if (k > 0 && br_bits(&r, k, &rem)) return fail("delta-golomb: truncated remainder");
        if (k > 0 && q > (UINT64_MAX >> k)) return fail("delta-golomb: delta overflows u64");
        uint64_t delta = (q << k) | rem;
        if (out[i - 1] > UINT64_MAX - delta) return fail("delta-golomb: indptr overflows u64");
        out[i] = out[i - 1] + delta;
    }
    return 9 + ((r.pos + 7) >> 3);
}

/* ------------------------------------------------------------------------
 * BitPacker4x (the bitpacking crate's 4-lane layout, as stated by the
 * reference's own device decoder, scx-gpu/kernels/forbp_decode_bp4x.cu:8-16):
 * 128 values, value p sits in lane p & 3 at slot p >> 2; lane l's stream is
 * the little-endian u32 sequence at bytes 16 w + 4 l (w = 0 .. bits - 1) and
 * slot j occupies bits [j * bits, (j + 1) * bits) of that stream.
 * ------------------------------------------------------------------------ */
void sho_bp4x_pack(const uint32_t *v, int32_t bits, uint8_t *out) {
    memset(out, 0, (size_t)bits * 16);
    for (int p = 0; p < 128; ++p) {
        int lane = p & 3, slot = p >> 2;
        uint64_t val = bits >= 32 ? v[p] : (v[p] & ((1u << bits) - 1));
        for (int b = 0; b < bits; ++b) {
            int64_t sbit = (int64_t)slot * bits + b; /* bit in the lane stream */
            int64_t w = sbit >> 5, bi = sbit & 31;
            if ((val >> b) & 1u) out[16 * w + 4 * lane + (bi >> 3)] |= (uint8_t)(1u << (bi & 7));
        }
    }
}

void sho_bp4x_unpack(const uint8_t *in, int32_t bits, uint32_t *v) {
    for (int p = 0; p < 128; ++p) {
        int lane = p & 3, slot = p >> 2;
        uint32_t val = 0;
        for (int b = 0; b < bits; ++b) {
            int64_t sbit = (int64_t)slot * bits + b;
            int64_t w = sbit >> 5, bi = sbit & 31;
            if ((in[16 * w + 4 * lane + (bi >> 3)] >> (bi & 7)) & 1u) val |= 1u << b;
        }
        v[p] = val;
    }
}

/* ------------------------------------------------------------------------
 * FOR-BP indices (scx-codec/src/forbp.rs:86-200 encode, :309-500 decode):
 * blocks of 128 rows; block header u32 block_nnz, u16 n_rows, LEB128 row nnz;
 * per non-empty row: frame_min (u16 or u32), frame_bits u8, then the deltas
 * (delta[0] = 0, delta[j] = idx[j] - idx[j-1]) at frame_bits each: rows of
 * >= 128 entries as BitPacker4x chunks of 128 plus a scalar LSB-first tail,
 * shorter rows scalar; each row's payload ends on a byte.
 * ------------------------------------------------------------------------ */
#define B_IDX 128
#define SIMD_THRESHOLD 128

static int bits_needed(uint32_t m) {
    int b = 0;
    while (m) {
        ++b;
        m >>= 1;
    }
    return b;
}

typedef struct {
    uint8_t *p;
    int64_t len, cap;
    int overflow;
} ob_t;

/* an output with p == NULL only counts bytes */
static uint8_t *ob_take(ob_t *o, int64_t n) {
    if (!o->p) {
        o->len += n;
        return NULL;
    }
    if (o->len + n > o->cap) {
        o->overflow = 1;
        o->len += n;
        return NULL;
    }
    uint8_t *r = o->p + o->len;
    o->len += n;
    return r;
}

static void write_varint(ob_t *o, uint32_t v) {
    do {
        uint8_t byte = v & 0x7F;
        v >>= 7;
        if (v) byte |= 0x80;
        uint8_t *q = ob_take(o, 1);
        if (q) *q = byte;
    } while (v);
}

/* read_varint (forbp.rs:42-66): the 5th byte may carry only 4 value bits */
static int read_varint(const uint8_t *in, int64_t len, int64_t *pos, uint32_t *v) {
    uint32_t res = 0;
    int shift = 0;
    for (;;) {
        if (*pos >= len) return -1;
        uint8_t byte = in[(*pos)++];
        if (shift == 28) {
            if (byte > 0x0F) return -1;
            *v = res | ((uint32_t)byte << 28);
            return 0;
        }
        res |= (uint32_t)(byte & 0x7F) << shift;
        if (!(byte & 0x80)) {
            *v = res;
            return 0;
        }
        shift += 7;
    }
}

int64_t sho_forbp_encode(const uint32_t *idx, const uint64_t *row_len, int64_t n_rows, int32_t u16, uint8_t *out,
                         int64_t cap) {
    ob_t o = {out, 0, cap, 0};
    int64_t off = 0;
    uint32_t *d = NULL;
    int64_t dcap = 0;
    for (int64_t b0 = 0; b0 < n_rows; b0 += B_IDX) {
        int64_t nb = n_rows - b0 < B_IDX ? n_rows - b0 : B_IDX;
        uint64_t bnnz = 0;
        for (int64_t r = 0; r < nb; ++r) bnnz += row_len[b0 + r];
        if (bnnz > 0xFFFFFFFFull) {
            free(d);
            return fail("forbp: block nnz overflows u32");
        }
        uint8_t *h = ob_take(&o, 6);
        if (h) {
            put_u32(h, (uint32_t)bnnz);
            put_u16(h + 4, (uint16_t)nb);
        }
        for (int64_t r = 0; r < nb; ++r) write_varint(&o, (uint32_t)row_len[b0 + r]);
        for (int64_t r = 0; r < nb; ++r) {
            int64_t nnz = (int64_t)row_len[b0 + r];
            if (nnz == 0) continue;
            const uint32_t *ri = idx + off;
            if (nnz > dcap) {
                dcap = nnz;
                d = realloc(d, sizeof(uint32_t) * (size_t)dcap);
            }
            d[0] = 0;
            uint32_t mx = 0;
            for (int64_t j = 1; j < nnz; ++j) {
                if (ri[j] < ri[j - 1]) {
                    free(d);
                    return fail("FOR-BP requires sorted row indices");
                }
                d[j] = ri[j] - ri[j - 1];
                if (d[j] > mx) mx = d[j];
            }
            int fb = bits_needed(mx);
            if (u16) {
                if (ri[0] > 0xFFFF) {
                    free(d);
                    return fail("forbp: frame_min above u16");
                }
                uint8_t *q = ob_take(&o, 2);
                if (q) put_u16(q, (uint16_t)ri[0]);
            } else {
                uint8_t *q = ob_take(&o, 4);
                if (q) put_u32(q, ri[0]);
            }
            uint8_t *q = ob_take(&o, 1);
            if (q) *q = (uint8_t)fb;
            if (fb > 0) {
                int64_t head = 0;
                if (nnz >= SIMD_THRESHOLD) {
                    int64_t full = nnz / SIMD_THRESHOLD;
                    for (int64_t c = 0; c < full; ++c) {
                        uint8_t *dst = ob_take(&o, (int64_t)fb * 16);
                        if (dst) sho_bp4x_pack(d + c * SIMD_THRESHOLD, fb, dst);
                    }
                    head = full * SIMD_THRESHOLD;
                }
                int64_t rem = nnz - head;
                if (rem > 0) {
                    int64_t nbytes = (rem * fb + 7) / 8;
                    uint8_t *dst = ob_take(&o, nbytes);
                    if (dst) {
                        bw_t w = {dst, nbytes, 0, 0};
                        for (int64_t j = head; j < nnz; ++j) bw_bits(&w, d[j], fb);
                        bw_pad(&w);
                    }
                }
            }
            off += nnz;
        }
    }
    free(d);
    if (o.overflow) return fail("forbp_encode: output capacity");
    return o.len;
}

int64_t sho_forbp_decode(const uint8_t *in, int64_t len, int64_t n_rows, int64_t nnz_hint, int32_t u16,
                         uint32_t *idx, int64_t idx_cap, uint64_t *row_len) {
    int64_t pos = 0, out = 0, rows_done = 0;
    uint32_t tmp[128];
    while (rows_done < n_rows) {
        if (pos + 6 > len) return fail("forbp: truncated block header");
        int64_t nb = get_u16(in + pos + 4);
        pos += 6;
        if (nb > n_rows - rows_done) return fail("forbp: block declares more rows than remain");
        uint32_t rn[65536];
        for (int64_t r = 0; r < nb; ++r)
            if (read_varint(in, len, &pos, &rn[r])) return fail("forbp: bad row nnz varint");
        for (int64_t r = 0; r < nb; ++r) {
            int64_t nnz = rn[r];
            row_len[rows_done + r] = (uint64_t)nnz;
            if (nnz == 0) continue;
            uint32_t fmin;
            if (u16) {
                if (pos + 2 > len) return fail("forbp: truncated frame_min");
                fmin = get_u16(in + pos);
                pos += 2;
            } else {
                if (pos + 4 > len) return fail("forbp: truncated frame_min");
                fmin = get_u32(in + pos);
                pos += 4;
            }
            if (pos + 1 > len) return fail("forbp: truncated frame_bits");
            int fb = in[pos++];
            if (fb > 32) return fail("forbp: frame_bits %d > 32", fb);
            if (out + nnz > idx_cap) return fail("forbp: decoded nnz exceeds the declared nnz");
            if (fb == 0) {
                for (int64_t j = 0; j < nnz; ++j) idx[out + j] = fmin;
                out += nnz;
                continue;
            }
            int64_t head = 0;
            if (nnz >= SIMD_THRESHOLD) {
                int64_t full = nnz / SIMD_THRESHOLD;
                for (int64_t c = 0; c < full; ++c) {
                    if (pos + (int64_t)fb * 16 > len) return fail("forbp: truncated BitPacker4x chunk");
                    sho_bp4x_unpack(in + pos, fb, tmp);
                    memcpy(idx + out + c * SIMD_THRESHOLD, tmp, sizeof tmp);
                    pos += (int64_t)fb * 16;
                }
                head = full * SIMD_THRESHOLD;
            }
            int64_t rem = nnz - head;
            if (rem > 0) {
                int64_t nbytes = (rem * fb + 7) / 8;
                if (pos + nbytes > len) return fail("forbp: truncated row payload");
                br_t br = {in + pos, nbytes, 0};
                for (int64_t j = head; j < nnz; ++j) {
                    uint64_t v;
                    br_bits(&br, fb, &v);
                    idx[out + j] = (uint32_t)v;
                }
                pos += nbytes;
            }
            /* prefix sum from frame_min, u32 overflow is an error (forbp.rs:478-483) */
            uint64_t prev = fmin;
            for (int64_t j = 0; j < nnz; ++j) {
                prev += idx[out + j];
                if (prev > 0xFFFFFFFFull) return fail("forbp: index overflows u32");
                idx[out + j] = (uint32_t)prev;
            }
            out += nnz;
        }
        rows_done += nb;
    }
    if (nnz_hint > 0 && out != nnz_hint) return fail("forbp: decoded %lld indices, declared %lld", (long long)out,
                                                     (long long)nnz_hint);
    return out;
}

/* ------------------------------------------------------------------------
 * Byte shuffle / unshuffle (scx-codec/src/shuffle.rs:18-79): byte b of
 * element i at b * n + i. Byte delta (scx-codec/src/byte_delta.rs): per plane
 * wrapping u8 differences, undone by a wrapping prefix sum.
 * ------------------------------------------------------------------------ */
void sho_byte_shuffle(const uint8_t *in, int64_t n, int32_t w, uint8_t *out) {
    for (int64_t i = 0; i < n; ++i)
        for (int b = 0; b < w; ++b) out[(int64_t)b * n + i] = in[i * w + b];
}

void sho_byte_unshuffle(const uint8_t *in, int64_t n, int32_t w, uint8_t *out) {
    for (int64_t i = 0; i < n; ++i)
        for (int b = 0; b < w; ++b) out[i * w + b] = in[(int64_t)b * n + i];
}

void sho_byte_delta(uint8_t *buf, int32_t planes, int64_t n) {
    if (n <= 1) return;
    for (int p = 0; p < planes; ++p)
        for (int64_t e = n - 1; e >= 1; --e) buf[p * n + e] = (uint8_t)(buf[p * n + e] - buf[p * n + e - 1]);
}

void sho_byte_undelta(uint8_t *buf, int32_t planes, int64_t n) {
    if (n <= 1) return;
    for (int p = 0; p < planes; ++p)
        for (int64_t e = 1; e < n; ++e) buf[p * n + e] = (uint8_t)(buf[p * n + e] + buf[p * n + e - 1]);
}

/* ------------------------------------------------------------------------
 * System zstd / LZ4 frame libraries, loaded at run time.
 * ------------------------------------------------------------------------ */
typedef size_t (*zstd_c_t)(void *, size_t, const void *, size_t, int);
typedef size_t (*zstd_d_t)(void *, size_t, const void *, size_t);
typedef size_t (*zstd_b_t)(size_t);
typedef unsigned (*zstd_e_t)(size_t);
typedef size_t (*lz4_cf_t)(void *, size_t, const void *, size_t, const void *);
typedef size_t (*lz4_cb_t)(size_t, const void *);
typedef size_t (*lz4_mk_t)(void **, unsigned);
typedef size_t (*lz4_fr_t)(void *);
typedef size_t (*lz4_d_t)(void *, void *, size_t *, const void *, size_t *, const void *);
typedef unsigned (*lz4_e_t)(size_t);

static struct {
    int tried;
    zstd_c_t c;
    zstd_d_t d;
    zstd_b_t bound;
    zstd_e_t err;
    lz4_cf_t lc;
    lz4_cb_t lb;
    lz4_mk_t lmk;
    lz4_fr_t lfr;
    lz4_d_t ld;
    lz4_e_t lerr;
} L;

static void load_libs(void) {
    if (L.tried) return;
    L.tried = 1;
    void *z = dlopen("libzstd.so.1", RTLD_NOW | RTLD_LOCAL);
    if (z) {
        L.c = (zstd_c_t)dlsym(z, "ZSTD_compress");
        L.d = (zstd_d_t)dlsym(z, "ZSTD_decompress");
        L.bound = (zstd_b_t)dlsym(z, "ZSTD_compressBound");
        L.err = (zstd_e_t)dlsym(z, "ZSTD_isError");
    }
    void *l = dlopen("liblz4.so.1", RTLD_NOW | RTLD_LOCAL);
    if (l) {
        L.lc = (lz4_cf_t)dlsym(l, "LZ4F_compressFrame");
        L.lb = (lz4_cb_t)dlsym(l, "LZ4F_compressFrameBound");
        L.lmk = (lz4_mk_t)dlsym(l, "LZ4F_createDecompressionContext");
        L.lfr = (lz4_fr_t)dlsym(l, "LZ4F_freeDecompressionContext");
        L.ld = (lz4_d_t)dlsym(l, "LZ4F_decompress");
        L.lerr = (lz4_e_t)dlsym(l, "LZ4F_isError");
    }
}

int32_t sho_have_zstd(void) {
    load_libs();
    return L.c && L.d && L.bound && L.err;
}

int32_t sho_have_lz4(void) {
    load_libs();
    return L.lc && L.lb && L.lmk && L.lfr && L.ld && L.lerr;
}

/* zstd level 3 (codecs/zstd_codec.rs:52-54, SHUFDELTA_ZSTD_LEVEL) */
static int64_t zstd_pack(const uint8_t *in, int64_t n, uint8_t **out) {
    if (!sho_have_zstd()) return fail("libzstd.so.1 not available");
    size_t cap = L.bound((size_t)n);
    *out = malloc(cap ? cap : 1);
    size_t r = L.c(*out, cap, in, (size_t)n, 3);
    if (L.err(r)) return fail("zstd compress failed");
    return (int64_t)r;
}

/* bounded decode: exactly `want` bytes or an error (zstd_decode_bounded +
 * expect_exact_len, codecs/zstd_codec.rs:63-80) */
static int zstd_unpack(const uint8_t *in, int64_t n, uint8_t *out, int64_t want) {
    if (!sho_have_zstd()) return fail("libzstd.so.1 not available");
    uint8_t dummy;
    size_t r = L.d(want ? out : &dummy, want ? (size_t)want : 1, in, (size_t)n);
    if (L.err(r)) return fail("zstd frame: decode error or more than %lld bytes", (long long)want);
    if ((int64_t)r != want) return fail("zstd frame: %zu bytes, expected %lld", r, (long long)want);
    return 0;
}

static int64_t lz4_pack(const uint8_t *in, int64_t n, uint8_t **out) {
    if (!sho_have_lz4()) return fail("liblz4.so.1 not available");
    size_t cap = L.lb((size_t)n, NULL);
    *out = malloc(cap);
    size_t r = L.lc(*out, cap, in, (size_t)n, NULL);
    if (L.lerr(r)) return fail("lz4 compress failed");
    return (int64_t)r;
}

static int lz4_unpack(const uint8_t *in, int64_t n, uint8_t *out, int64_t want) {
    if (!sho_have_lz4()) return fail("liblz4.so.1 not available");
    void *ctx = NULL;
    if (L.lerr(L.lmk(&ctx, 100))) return fail("lz4 context");
    int64_t got = 0, used = 0;
    int rc = 0;
    uint8_t spill[64];
    for (;;) {
        size_t dst = want - got > 0 ? (size_t)(want - got) : sizeof spill;
        uint8_t *dp = want - got > 0 ? out + got : spill;
        size_t src = (size_t)(n - used);
        size_t r = L.ld(ctx, dp, &dst, in + used, &src, NULL);
        if (L.lerr(r)) {
            rc = fail("lz4 frame: decode error");
            break;
        }
        if (dp == spill && dst > 0) {
            rc = fail("lz4 frame: more than %lld bytes", (long long)want);
            break;
        }
        got += (int64_t)dst;
        used += (int64_t)src;
        if (r == 0) break;
        if (src == 0 && dst == 0) {
            rc = fail("lz4 frame: truncated");
            break;
        }
    }
    L.lfr(ctx);
    if (!rc && got != want) rc = fail("lz4 frame: %lld bytes, expected %lld", (long long)got, (long long)want);
    return rc;
}

/* ------------------------------------------------------------------------
 * Shards (scx-format/src/shard.rs:98-178 header, :410-436 block index entry,
 * :594-730 resolve_block_index; scx-codec/src/dispatch.rs:201-260
 * decode_shard_scipy; scx-format-io/src/encoder.rs:524-690 framing).
 * ------------------------------------------------------------------------ */
enum { C_NONE = 0, C_SCX1 = 1, C_ZSTD = 2, C_LZ4 = 3, C_PCODEC = 4, C_SHUF = 5 };
enum { V_U8 = 0, V_U16 = 1, V_U32 = 2, V_F32 = 3, V_F16 = 4 };
#define HDR 76
#define BIE 22

static int venc_width(int v) { return v == V_U8 ? 1 : (v == V_U16 || v == V_F16) ? 2 : 4; }
static int venc_integer(int v) { return v <= V_U32; }

typedef struct {
    uint8_t *ip, *ix, *vv;
    int64_t nip, nix, nvv;
} streams_t;

static void free_streams(streams_t *s) {
    free(s->ip);
    free(s->ix);
    free(s->vv);
    memset(s, 0, sizeof *s);
}

/* encode_shard for one (sub-)shard: codecs None / Scx1 / Zstd / Lz4Shuffle /
 * ShufDeltaZstd (codecs/{none,scx1,zstd_codec,lz4_shuffle,shufdelta}.rs) */
static int encode_streams(const uint64_t *indptr, int64_t n_rows, const uint32_t *idx, const uint8_t *vals,
                          int codec, int venc, int u16, streams_t *s) {
    memset(s, 0, sizeof *s);
    int64_t nnz = (int64_t)(indptr[n_rows] - indptr[0]);
    int iw = u16 ? 2 : 4, vw = venc_width(venc);
    uint8_t *ipraw = malloc((size_t)(n_rows + 1) * 8 + 1);
    for (int64_t i = 0; i <= n_rows; ++i) put_u64(ipraw + 8 * i, indptr[i]);
    uint8_t *ixraw = malloc((size_t)nnz * iw + 1);
    for (int64_t i = 0; i < nnz; ++i) {
        if (u16) {
            if (idx[i] > 0xFFFF) {
                free(ipraw);
                free(ixraw);
                return fail("index above u16");
            }
            put_u16(ixraw + 2 * i, (uint16_t)idx[i]);
        } else
            put_u32(ixraw + 4 * i, idx[i]);
    }
    int rc = 0;
    if (codec == C_NONE) {
        s->ip = ipraw, s->nip = (n_rows + 1) * 8;
        s->ix = ixraw, s->nix = nnz * iw;
        s->vv = malloc((size_t)nnz * vw + 1), s->nvv = nnz * vw;
        memcpy(s->vv, vals, (size_t)nnz * vw);
        return 0;
    }
    if (codec == C_SCX1) {
        if (!venc_integer(venc)) {
            rc = fail("Scx1 codec does not support float value encodings");
            goto out;
        }
        int64_t cap = sho_dg_encode(indptr, n_rows + 1, NULL, 0);
        s->ip = malloc((size_t)(cap > 0 ? cap : 1));
        s->nip = cap < 0 ? -1 : sho_dg_encode(indptr, n_rows + 1, s->ip, cap);
        uint64_t *rl = malloc(sizeof(uint64_t) * (size_t)(n_rows + 1));
        for (int64_t r = 0; r < n_rows; ++r) rl[r] = indptr[r + 1] - indptr[r];
        cap = sho_forbp_encode(idx, rl, n_rows, u16, NULL, 0);
        s->ix = malloc((size_t)(cap > 0 ? cap : 1));
        s->nix = cap < 0 ? -1 : sho_forbp_encode(idx, rl, n_rows, u16, s->ix, cap);
        free(rl);
        uint32_t *v32 = malloc(sizeof(uint32_t) * (size_t)(nnz + 1));
        for (int64_t i = 0; i < nnz; ++i)
            v32[i] = vw == 1 ? vals[i] : vw == 2 ? get_u16(vals + 2 * i) : get_u32(vals + 4 * i);
        cap = sho_rice_encode(v32, nnz, NULL, 0);
        s->vv = malloc((size_t)(cap > 0 ? cap : 1));
        s->nvv = cap < 0 ? -1 : sho_rice_encode(v32, nnz, s->vv, cap);
        free(v32);
        if (s->nip < 0 || s->nix < 0 || s->nvv < 0) rc = -1;
        goto out;
    }
    if (codec == C_ZSTD) {
        s->nip = zstd_pack(ipraw, (n_rows + 1) * 8, &s->ip);
        s->nix = zstd_pack(ixraw, nnz * iw, &s->ix);
        s->nvv = zstd_pack(vals, nnz * vw, &s->vv);
        if (s->nip < 0 || s->nix < 0 || s->nvv < 0) rc = -1;
        goto out;
    }
    if (codec == C_LZ4 || codec == C_SHUF) {
        uint8_t *a = malloc((size_t)(n_rows + 1) * 8 + 1), *b = malloc((size_t)nnz * iw + 1),
                *c = malloc((size_t)nnz * vw + 1);
        sho_byte_shuffle(ipraw, n_rows + 1, 8, a);
        sho_byte_shuffle(ixraw, nnz, iw, b);
        int shuffle_vals = codec == C_LZ4 || venc_integer(venc);
        if (shuffle_vals)
            sho_byte_shuffle(vals, nnz, vw, c);
        else
            memcpy(c, vals, (size_t)nnz * vw);
        if (codec == C_LZ4) {
            s->nip = lz4_pack(a, (n_rows + 1) * 8, &s->ip);
            s->nix = lz4_pack(b, nnz * iw, &s->ix);
            s->nvv = lz4_pack(c, nnz * vw, &s->vv);
        } else {
            sho_byte_delta(a, 8, n_rows + 1);
            sho_byte_delta(b, iw, nnz);
            s->nip = zstd_pack(a, (n_rows + 1) * 8, &s->ip);
            s->nix = zstd_pack(b, nnz * iw, &s->ix);
            s->nvv = zstd_pack(c, nnz * vw, &s->vv);
        }
        free(a);
        free(b);
        free(c);
        if (s->nip < 0 || s->nix < 0 || s->nvv < 0) rc = -1;
        goto out;
    }
    rc = fail("no encoder for codec %d in the oracle", codec);
out:
    free(ipraw);
    free(ixraw);
    if (rc) free_streams(s);
    return rc;
}

int64_t sho_encode_shard(const uint64_t *indptr, int64_t n_rows, const uint32_t *idx, const uint8_t *values_raw,
                         int32_t codec, int32_t venc, int32_t u16, uint32_t n_minor, uint64_t global_offset,
                         int32_t row_group_rows, uint8_t *out, int64_t cap) {
    int framed = row_group_rows > 0;
    int vw = venc_width(venc);
    if (framed && n_rows == 0) return fail("a framed shard needs at least one row");
    int64_t g = framed ? row_group_rows : (n_rows > 0 ? n_rows : 1);
    if (g > 65535) g = 65535;
    int64_t ngroups = framed ? (n_rows + g - 1) / g : 1;
    streams_t *gs = calloc((size_t)ngroups, sizeof(streams_t));
    uint32_t *ipo = calloc((size_t)ngroups + 1, 4), *ixo = calloc((size_t)ngroups + 1, 4),
             *vvo = calloc((size_t)ngroups + 1, 4);
    int64_t rc = 0;
    uint64_t *loc = malloc(sizeof(uint64_t) * (size_t)(g + 2));
    for (int64_t gi = 0; gi < ngroups; ++gi) {
        int64_t r0 = framed ? gi * g : 0, r1 = framed ? (r0 + g < n_rows ? r0 + g : n_rows) : n_rows;
        uint64_t base = framed ? indptr[r0] : 0;
        const uint64_t *ip = indptr + r0;
        if (framed) {
            for (int64_t r = r0; r <= r1; ++r) loc[r - r0] = indptr[r] - base;
            ip = loc;
        }
        if (encode_streams(ip, r1 - r0, idx + base, values_raw + base * vw, codec, venc, u16, &gs[gi])) {
            rc = -1;
            break;
        }
        ipo[gi + 1] = ipo[gi] + (uint32_t)gs[gi].nip;
        ixo[gi + 1] = ixo[gi] + (uint32_t)gs[gi].nix;
        vvo[gi + 1] = vvo[gi] + (uint32_t)gs[gi].nvv;
    }
    free(loc);
    int64_t total = 0;
    if (!rc) {
        int64_t nent = framed ? ngroups : (n_rows > 65535 ? (n_rows + 65534) / 65535 : 1);
        int64_t bil = 4 + BIE * nent;
        total = HDR + ipo[ngroups] + ixo[ngroups] + vvo[ngroups] + bil;
        if (out) {
            if (total > cap) {
                rc = fail("encode_shard: output capacity");
            } else {
                memset(out, 0, HDR);
                memcpy(out, "SCXS", 4);
                out[4] = (uint8_t)(framed ? 2 : 1);
                out[5] = 0;
                out[6] = (uint8_t)codec;
                out[7] = (uint8_t)venc;
                out[8] = (uint8_t)(u16 ? 0 : 1);
                put_u32(out + 12, (uint32_t)n_rows);
                put_u32(out + 16, n_minor);
                uint64_t nnz = n_rows ? indptr[n_rows] - indptr[0] : 0;
                put_u64(out + 20, nnz);
                put_u64(out + 28, global_offset);
                uint32_t o_ip = HDR, o_ix = o_ip + ipo[ngroups], o_vv = o_ix + ixo[ngroups],
                         o_bi = o_vv + vvo[ngroups];
                put_u32(out + 36, o_ip);
                put_u32(out + 40, ipo[ngroups]);
                put_u32(out + 44, o_ix);
                put_u32(out + 48, ixo[ngroups]);
                put_u32(out + 52, o_vv);
                put_u32(out + 56, vvo[ngroups]);
                put_u32(out + 60, o_bi);
                put_u32(out + 64, (uint32_t)bil);
                for (int64_t gi = 0; gi < ngroups; ++gi) {
                    memcpy(out + o_ip + ipo[gi], gs[gi].ip, (size_t)gs[gi].nip);
                    memcpy(out + o_ix + ixo[gi], gs[gi].ix, (size_t)gs[gi].nix);
                    memcpy(out + o_vv + vvo[gi], gs[gi].vv, (size_t)gs[gi].nvv);
                }
                uint8_t *bi = out + o_bi;
                put_u32(bi, (uint32_t)nent);
                int64_t step = framed ? g : 65535;
                for (int64_t gi = 0; gi < nent; ++gi) {
                    uint8_t *e = bi + 4 + BIE * gi;
                    int64_t r0 = gi * step, r1 = r0 + step < n_rows ? r0 + step : n_rows;
                    put_u32(e, (uint32_t)r0);
                    put_u16(e + 4, (uint16_t)(r1 - r0));
                    put_u32(e + 6, framed ? ipo[gi] : 0);
                    put_u32(e + 10, framed ? ixo[gi] : 0);
                    put_u32(e + 14, framed ? vvo[gi] : 0);
                    put_u32(e + 18, (uint32_t)(n_rows ? indptr[r1] - indptr[r0] : 0));
                }
            }
        }
    }
    for (int64_t gi = 0; gi < ngroups; ++gi) free_streams(&gs[gi]);
    free(gs);
    free(ipo);
    free(ixo);
    free(vvo);
    return rc ? -1 : total;
}

typedef struct {
    int version, codec, venc, u16;
    int64_t n_major, n_minor, nnz;
    uint32_t o[8]; /* ip off/len, ix off/len, vv off/len, bi off/len */
} hdr_t;

static int read_hdr(const uint8_t *sh, int64_t len, hdr_t *h) {
    if (len < HDR) return fail("shard too small: %lld bytes < 76 header", (long long)len);
    if (memcmp(sh, "SCXS", 4)) return fail("shard header: bad magic");
    h->version = sh[4];
    if (h->version > 2) return fail("shard header: unsupported shard format version %d", h->version);
    h->codec = sh[6];
    h->venc = sh[7];
    h->u16 = sh[8] == 0;
    h->n_major = get_u32(sh + 12);
    h->n_minor = get_u32(sh + 16);
    h->nnz = (int64_t)get_u64(sh + 20);
    for (int i = 0; i < 8; ++i) h->o[i] = get_u32(sh + 36 + 4 * i);
    for (int i = 0; i < 4; ++i)
        if ((int64_t)h->o[2 * i] + h->o[2 * i + 1] > len) return fail("shard: sub-stream %d exceeds the shard", i);
    if (h->codec > C_SHUF) return fail("unknown codec_id: %d", h->codec);
    if (h->venc > V_F16) return fail("unknown value_encoding: %d", h->venc);
    return 0;
}

int32_t sho_shard_info(const uint8_t *sh, int64_t len, int64_t *n_major, int64_t *n_minor, int64_t *nnz,
                       int32_t *codec, int32_t *venc, int32_t *index_dtype, int32_t *version) {
    hdr_t h;
    if (read_hdr(sh, len, &h)) return -1;
    *n_major = h.n_major, *n_minor = h.n_minor, *nnz = h.nnz, *codec = h.codec, *venc = h.venc;
    *index_dtype = h.u16 ? 0 : 1, *version = h.version;
    return 0;
}

static float half_to_float(uint16_t hb) {
    int s = hb >> 15, e = (hb >> 10) & 31, m = hb & 1023;
    float v;
    if (e == 0)
        v = ldexpf((float)m, -24);
    else if (e == 31)
        v = m ? NAN : INFINITY;
    else
        v = ldexpf((float)(m | 1024), e - 25);
    return s ? -v : v;
}

static float raw_value(const uint8_t *p, int venc) {
    switch (venc) {
    case V_U8: return (float)p[0];
    case V_U16: return (float)get_u16(p);
    case V_U32: return (float)get_u32(p);
    case V_F32: {
        uint32_t b = get_u32(p);
        float f;
        memcpy(&f, &b, 4);
        return f;
    }
    default: return half_to_float(get_u16(p));
    }
}

/* one (sub-)shard: local indptr (n_rows + 1, from 0), indices, values */
static int decode_streams(const hdr_t *h, const uint8_t *ip, int64_t nip, const uint8_t *ix, int64_t nix,
                          const uint8_t *vv, int64_t nvv, int64_t n_rows, int64_t nnz, uint64_t *indptr,
                          uint32_t *idx, float *data) {
    int iw = h->u16 ? 2 : 4, vw = venc_width(h->venc);
    int codec = h->codec;
    if (codec == C_SCX1) {
        if (!venc_integer(h->venc)) return fail("Scx1 codec does not support float value encodings");
        if (n_rows + 1 > nip * 8 || nnz > nix * 8 || nnz > nvv * 8)
            return fail("scx1: declared sizes exceed the sub-streams' capacity");
        if (sho_dg_decode(ip, nip, n_rows + 1, indptr) < 0) return -1;
        uint64_t *rl = malloc(sizeof(uint64_t) * (size_t)(n_rows + 1));
        int64_t got = sho_forbp_decode(ix, nix, n_rows, nnz, h->u16, idx, nnz, rl);
        free(rl);
        if (got < 0) return -1;
        if (got != nnz) return fail("forbp: decoded %lld indices, declared %lld", (long long)got, (long long)nnz);
        uint32_t *v32 = malloc(sizeof(uint32_t) * (size_t)(nnz + 1));
        if (sho_rice_decode(vv, nvv, nnz, v32) < 0) {
            free(v32);
            return -1;
        }
        for (int64_t i = 0; i < nnz; ++i) data[i] = (float)v32[i];
        free(v32);
        return 0;
    }
    int64_t nb_ip = (n_rows + 1) * 8, nb_ix = nnz * iw, nb_vv = nnz * vw;
    uint8_t *a = malloc((size_t)nb_ip + 1), *b = malloc((size_t)nb_ix + 1), *c = malloc((size_t)nb_vv + 1);
    int rc = 0;
    if (codec == C_NONE) {
        if (nip != nb_ip || nix != nb_ix || nvv != nb_vv) rc = fail("none codec: sub-stream lengths disagree");
        else {
            memcpy(a, ip, (size_t)nb_ip);
            memcpy(b, ix, (size_t)nb_ix);
            memcpy(c, vv, (size_t)nb_vv);
        }
    } else if (codec == C_ZSTD) {
        rc = zstd_unpack(ip, nip, a, nb_ip) || zstd_unpack(ix, nix, b, nb_ix) || zstd_unpack(vv, nvv, c, nb_vv);
    } else if (codec == C_LZ4 || codec == C_SHUF) {
        uint8_t *pa = malloc((size_t)nb_ip + 1), *pb = malloc((size_t)nb_ix + 1), *pc = malloc((size_t)nb_vv + 1);
        if (codec == C_LZ4)
            rc = lz4_unpack(ip, nip, pa, nb_ip) || lz4_unpack(ix, nix, pb, nb_ix) || lz4_unpack(vv, nvv, pc, nb_vv);
        else
            rc = zstd_unpack(ip, nip, pa, nb_ip) || zstd_unpack(ix, nix, pb, nb_ix) ||
                 zstd_unpack(vv, nvv, pc, nb_vv);
        if (!rc) {
            if (codec == C_SHUF) {
                sho_byte_undelta(pa, 8, n_rows + 1);
                sho_byte_undelta(pb, iw, nnz);
            }
            sho_byte_unshuffle(pa, n_rows + 1, 8, a);
            sho_byte_unshuffle(pb, nnz, iw, b);
            if (codec == C_LZ4 || venc_integer(h->venc))
                sho_byte_unshuffle(pc, nnz, vw, c);
            else
                memcpy(c, pc, (size_t)nb_vv);
        }
        free(pa);
        free(pb);
        free(pc);
    } else {
        rc = fail("no decoder for codec %d in the oracle", codec);
    }
    if (!rc) {
        for (int64_t i = 0; i <= n_rows; ++i) indptr[i] = get_u64(a + 8 * i);
        for (int64_t i = 0; i < nnz; ++i) idx[i] = h->u16 ? get_u16(b + 2 * i) : get_u32(b + 4 * i);
        for (int64_t i = 0; i < nnz; ++i) data[i] = raw_value(c + (int64_t)vw * i, h->venc);
    }
    free(a);
    free(b);
    free(c);
    return rc;
}

/* check_indptr_shape (scx-codec/src/guards.rs:258-305) */
static int check_indptr(const uint64_t *ip, int64_t n_rows, int64_t nnz) {
    if (ip[0] != 0) return fail("decoded CSR indptr must start at 0, got %llu", (unsigned long long)ip[0]);
    for (int64_t r = 1; r <= n_rows; ++r)
        if (ip[r] < ip[r - 1]) return fail("decoded CSR indptr not monotone at row %lld", (long long)(r - 1));
    if ((int64_t)ip[n_rows] != nnz)
        return fail("decoded CSR indptr ends at %llu, nnz is %lld", (unsigned long long)ip[n_rows], (long long)nnz);
    return 0;
}

int32_t sho_decode_shard(const uint8_t *sh, int64_t len, int64_t *indptr, int32_t *idx_out, float *data) {
    hdr_t h;
    if (read_hdr(sh, len, &h)) return -1;
    if (h.codec == C_PCODEC) return fail("no decoder for codec pcodec in the oracle");
    const uint8_t *ip = sh + h.o[0], *ix = sh + h.o[2], *vv = sh + h.o[4], *bi = sh + h.o[6];
    int64_t n = h.n_major, nnz = h.nnz;
    uint64_t *gip = malloc(sizeof(uint64_t) * (size_t)(n + 2));
    uint32_t *idx = malloc(sizeof(uint32_t) * (size_t)(nnz + 1));
    int rc = 0;
    if (h.version < 2) {
        rc = decode_streams(&h, ip, h.o[1], ix, h.o[3], vv, h.o[5], n, nnz, gip, idx, data);
        if (!rc) rc = check_indptr(gip, n, nnz);
    } else {
        /* resolve_block_index (shard.rs:594-730) */
        int64_t bl = h.o[7];
        if (bl < 4) rc = fail("block index: truncated");
        int64_t ng = rc ? 0 : get_u32(bi);
        if (!rc && 4 + ng * BIE > bl) rc = fail("block index: %lld entries exceed its length", (long long)ng);
        if (!rc && ng == 0 && !(n == 0 && nnz == 0))
            rc = fail("empty block index, but the header declares n_major %lld / nnz %lld", (long long)n,
                      (long long)nnz);
        int64_t row = 0, nsum = 0;
        gip[0] = 0;
        for (int64_t gi = 0; gi < ng && !rc; ++gi) {
            const uint8_t *e = bi + 4 + BIE * gi;
            int64_t rs = get_u32(e), nr = get_u16(e + 4), gnnz = get_u32(e + 18);
            uint32_t o0[3] = {get_u32(e + 6), get_u32(e + 10), get_u32(e + 14)};
            uint32_t o1[3];
            if (gi + 1 < ng) {
                const uint8_t *e2 = e + BIE;
                o1[0] = get_u32(e2 + 6), o1[1] = get_u32(e2 + 10), o1[2] = get_u32(e2 + 14);
            } else {
                o1[0] = h.o[1], o1[1] = h.o[3], o1[2] = h.o[5];
            }
            if (rs != row) rc = fail("entry %lld: row_start %lld != expected %lld", (long long)gi, (long long)rs,
                                     (long long)row);
            else if (nr == 0) rc = fail("entry %lld: n_rows == 0", (long long)gi);
            else if (gnnz > nr * h.n_minor) rc = fail("entry %lld: nnz_in_block exceeds group capacity", (long long)gi);
            else if (gi == 0 && (o0[0] | o0[1] | o0[2])) rc = fail("first entry sub-stream offsets must all be 0");
            for (int s = 0; s < 3 && !rc; ++s) {
                uint32_t lim = s == 0 ? h.o[1] : s == 1 ? h.o[3] : h.o[5];
                if (o0[s] > o1[s] || o1[s] > lim) rc = fail("entry %lld: sub-stream range invalid", (long long)gi);
            }
            if (!rc && o1[0] == o0[0]) rc = fail("entry %lld: empty indptr range", (long long)gi);
            if (!rc && gnnz > 0 && (o1[1] == o0[1] || o1[2] == o0[2]))
                rc = fail("entry %lld: empty indices/values range with nnz > 0", (long long)gi);
            if (!rc && nsum + gnnz > nnz) rc = fail("block index: group nnz exceed the header's nnz");
            if (rc) break;
            uint64_t *lip = malloc(sizeof(uint64_t) * (size_t)(nr + 1));
            rc = decode_streams(&h, ip + o0[0], o1[0] - o0[0], ix + o0[1], o1[1] - o0[1], vv + o0[2], o1[2] - o0[2],
                                nr, gnnz, lip, idx + nsum, data + nsum);
            if (!rc) rc = check_indptr(lip, nr, gnnz);
            if (!rc)
                for (int64_t r = 1; r <= nr; ++r) gip[row + r] = (uint64_t)nsum + lip[r];
            free(lip);
            row += nr;
            nsum += gnnz;
        }
        if (!rc && row != n) rc = fail("coverage %lld != header.n_major %lld", (long long)row, (long long)n);
        if (!rc && nsum != nnz) rc = fail("block index nnz %lld != header.nnz %lld", (long long)nsum, (long long)nnz);
    }
    if (!rc) {
        /* u32 -> i32 with the index bound (dispatch.rs:250 u32_vec_to_i32_bounded) */
        for (int64_t i = 0; i < nnz && !rc; ++i) {
            if (idx[i] > 0x7FFFFFFFu) rc = fail("index %u does not fit i32", idx[i]);
            else if (h.n_minor > 0 && idx[i] >= (uint64_t)h.n_minor)
                rc = fail("index %u out of range for n_minor %lld", idx[i], (long long)h.n_minor);
            else
                idx_out[i] = (int32_t)idx[i];
        }
        for (int64_t r = 0; r <= n && !rc; ++r) indptr[r] = (int64_t)gip[r];
    }
    free(gip);
    free(idx);
    return rc;
}
