/* CPU oracle for the SCX shard codecs and the shard container -- test
 * infrastructure only (tests/, smoke() and bench.py's cpu_baseline leg).
 *
 * Restates, in plain C, the scalar reference decoders and encoders of
 * scx-codec (Rice values, Delta-Golomb indptr, FOR-BP indices with the
 * BitPacker4x layout for rows of >= 128 entries, byte shuffle / byte delta)
 * and the 76-byte shard header + block index of scx-format, so the GPU
 * decoder can be checked against an independent implementation. zstd and LZ4
 * frames go through the system libzstd.so.1 / liblz4.so.1 (dlopen). */
#ifndef SCXSHARD_ORACLE_H
#define SCXSHARD_ORACLE_H
#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- codecs on flat buffers (return bytes written, or -1 on error) ---- */
int64_t sho_rice_encode(const uint32_t *v, int64_t n, uint8_t *out, int64_t cap);
int64_t sho_rice_decode(const uint8_t *in, int64_t len, int64_t n, uint32_t *out);
int64_t sho_dg_encode(const uint64_t *indptr, int64_t n, uint8_t *out, int64_t cap);
int64_t sho_dg_decode(const uint8_t *in, int64_t len, int64_t n, uint64_t *out);
int64_t sho_forbp_encode(const uint32_t *idx, const uint64_t *row_len, int64_t n_rows, int32_t u16,
                         uint8_t *out, int64_t cap);
/* returns nnz decoded (or -1); row_len receives n_rows entries */
int64_t sho_forbp_decode(const uint8_t *in, int64_t len, int64_t n_rows, int64_t nnz_hint, int32_t u16,
                         uint32_t *idx, int64_t idx_cap, uint64_t *row_len);
void sho_bp4x_pack(const uint32_t *v128, int32_t bits, uint8_t *out);
void sho_bp4x_unpack(const uint8_t *in, int32_t bits, uint32_t *v128);
void sho_byte_shuffle(const uint8_t *in, int64_t n, int32_t w, uint8_t *out);
void sho_byte_unshuffle(const uint8_t *in, int64_t n, int32_t w, uint8_t *out);
void sho_byte_delta(uint8_t *buf, int32_t planes, int64_t n);
void sho_byte_undelta(uint8_t *buf, int32_t planes, int64_t n);

/* ---- system compressors (0 when the library is absent) ---- */
int32_t sho_have_zstd(void);
int32_t sho_have_lz4(void);

/* ---- whole shards ---- */
/* Encodes one CSR shard section (76-byte header, the three sub-streams, block
 * index). row_group_rows > 0 frames it (shard v2). values_raw holds nnz values
 * in the value encoding's width. Returns the section size, or -1 (the error
 * text is in sho_last_error()). Pass out=NULL to learn the size. */
int64_t sho_encode_shard(const uint64_t *indptr, int64_t n_rows, const uint32_t *idx, const uint8_t *values_raw,
                         int32_t codec, int32_t value_encoding, int32_t u16, uint32_t n_minor, uint64_t global_offset,
                         int32_t row_group_rows, uint8_t *out, int64_t cap);
/* Reads the header fields: n_major, n_minor, nnz, codec, value encoding,
 * index dtype, shard format version. 0 on success. */
int32_t sho_shard_info(const uint8_t *sh, int64_t len, int64_t *n_major, int64_t *n_minor, int64_t *nnz,
                       int32_t *codec, int32_t *venc, int32_t *index_dtype, int32_t *version);
/* decode_shard_scipy: i64 indptr, i32 indices, f32 data. 0 on success. */
int32_t sho_decode_shard(const uint8_t *sh, int64_t len, int64_t *indptr, int32_t *idx, float *data);
const char *sho_last_error(void);

#ifdef __cplusplus
}
#endif
#endif
