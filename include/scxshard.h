/*
 * scxshard — on-GPU decode of SCX CSR shards into a device-resident CSR (C ABI).
 *
 * Replaces the reference's device decode entry points
 *   scx_gpu::decode_shard_gpu(dev, shard_bytes) -> GpuCsr
 *       (scx-gpu/src/shard_decode.rs:251-285)
 *   scx_gpu::decode_csr_shards_to_device_with_stats(dev, shards)
 *       -> (GpuCsr, DeviceDecodeStats)   (scx-gpu/src/gpu_csr_assemble.rs:47-140)
 * whose GpuCsr is i64 indptr / i32 indices / f32 data
 * (scx-gpu/src/shard_decode.rs:65-74). Here the caller owns the three device
 * arrays: scxshard_read_info gives each shard's rows and nnz, the caller sums
 * them and allocates (n_rows + 1, nnz, nnz) elements, and scxshard_decode_csr
 * fills them. The result can be handed straight to scxpca_matrix_create.
 *
 * Codecs (shard header codec_id): Scx1 (Delta-Golomb indptr on the host,
 * FOR-BP indices -- scalar and BitPacker4x rows -- and Rice values on the
 * device), None (device widening), Zstd / Lz4Shuffle / ShufDeltaZstd (frames
 * decompressed on the host through the system libzstd.so.1 / liblz4.so.1, the
 * byte un-delta, un-shuffle and widening on the device). Pcodec shards are
 * refused with SCXPCA_ERR_INVALID. Unframed (shard v1) and row-group-framed
 * (shard v2) shards are both accepted.
 *
 * Errors: a malformed shard (header, block index, sub-stream, an index at or
 * past n_minor or above 2^31 - 1, an indptr that is not a CSR row pointer)
 * returns SCXPCA_ERR_INVALID with the reason in scxpca_last_error; the output
 * arrays are then unspecified. The call returns after the device work ends.
 */
#ifndef SCXSHARD_H
#define SCXSHARD_H

#include <stddef.h>
#include <stdint.h>

#include "scxpca.h"

#ifdef __cplusplus
extern "C" {
#endif

/* The 76-byte shard header (scx-format/src/shard.rs:36-94), read on the host. */
typedef struct {
    int64_t n_major;             /* rows */
    int64_t n_minor;             /* columns (0 = not declared) */
    int64_t nnz;
    int64_t global_offset;       /* first row in the full matrix */
    int32_t codec_id;            /* 0 none, 1 scx1, 2 zstd, 3 lz4shuffle, 4 pcodec, 5 shufdelta */
    int32_t value_encoding;      /* 0 u8, 1 u16, 2 u32, 3 f32, 4 f16 */
    int32_t index_dtype;         /* 0 u16, 1 u32 */
    int32_t shard_format_version;/* 1 whole shard, 2 row-group framed */
    int32_t shard_type;          /* 0 CSR, 1 CSC */
    int32_t n_groups;            /* block-index entries of a framed shard, else 1 */
} scxshard_info;

/* DeviceDecodeStats (scx-gpu/src/shard_decode.rs:186-240) plus timings. */
typedef struct {
    uint64_t host_uploaded_bytes;   /* every byte copied host -> device */
    uint64_t device_decoded_bytes;  /* index + value bytes produced by device kernels */
    int32_t fully_device_decoded;   /* 1 when every shard's indices and values were
                                       decoded on the device from their encoded bytes
                                       (Scx1 / None) */
    uint32_t n_shards_scx1_gpu;
    uint32_t n_shards_shufdelta_gpu;
    uint32_t n_shards_host_bounced; /* always 0 here: no codec is decoded wholly on the host */
    uint32_t n_shards_other_gpu;    /* None / Zstd / Lz4Shuffle shards */
    uint32_t n_groups;              /* row groups decoded (1 per unframed shard) */
    double host_seconds;            /* header parse, indptr, prescans, decompression */
    double upload_seconds;          /* host -> device copies */
    double device_seconds;          /* decode kernels (HIP events) */
    double total_seconds;
} scxshard_stats;

/* Parses and checks the header (and a framed shard's block index). */
scxpca_status scxshard_read_info(const uint8_t *shard, uint64_t len, scxshard_info *out);

/* Decodes n_shards shards, row-stacked in the given order, into one CSR:
 * indptr[n_rows + 1] (i64), indices[nnz] (i32), data[nnz] (f32), all device
 * pointers on `device`. n_rows and nnz must equal the sums over the shard
 * headers, and every shard must declare the same n_minor
 * (scx-gpu/src/gpu_csr_assemble.rs:70-90). stream: a hipStream_t on `device`
 * or NULL (the library's own). stats may be NULL. */
scxpca_status scxshard_decode_csr(const uint8_t *const *shards, const uint64_t *lens, int64_t n_shards,
                                  int64_t *indptr, int32_t *indices, float *data, int64_t n_rows, int64_t nnz,
                                  int32_t device, void *stream, scxshard_stats *stats);

/* Writes rows [0, n_rows) of a host CSR as one CSR shard section -- the
 * layout of the reference writer's encode_shard / encode_shard_framed
 * (scx-codec/src/dispatch.rs:36-60, scx-format-io/src/encoder.rs:524-690,
 * header and block index scx-format/src/shard.rs:98-122, :440-478) -- for
 * codec 1 (Scx1; integer values) or 0 (None). values: n = indptr[n_rows] -
 * indptr[0] elements of the value encoding's width (0 u8, 1 u16, 2 u32, 3 f32,
 * 4 f16). index_dtype 0 writes u16 indices. row_group_rows > 0 frames the
 * shard (shard format 2, one block-index entry per group). out == NULL
 * returns the size in *written; otherwise cap must hold it. The 8 checksum
 * bytes of the header are left zero (no reader verifies them). Host only. */
scxpca_status scxshard_encode(const int64_t *indptr, int64_t n_rows, const uint32_t *indices, const void *values,
                              int32_t value_encoding, int32_t codec, int32_t index_dtype, int64_t n_minor,
                              int64_t global_offset, int32_t row_group_rows, uint8_t *out, uint64_t cap,
                              uint64_t *written);

#ifdef __cplusplus
}
#endif
#endif
