"""SCX shards onto the GPU: the host mirror of the reference's device decode.

``decode_shard_gpu`` and ``decode_csr_shards_to_device`` mirror
``scx_gpu::decode_shard_gpu`` (scx-gpu/src/shard_decode.rs:251-285) and
``scx_gpu::decode_csr_shards_to_device_with_stats``
(scx-gpu/src/gpu_csr_assemble.rs:47-140): they take raw shard sections
(76-byte header + encoded sub-streams + block index) and return a
device-resident CSR -- i64 indptr, i32 indices, f32 data, the reference's
``GpuCsr`` layout (scx-gpu/src/shard_decode.rs:65-74) -- plus the transfer
statistics (``DeviceDecodeStats``, :186-240). The decoding runs in
``libscxpca.so`` (include/scxshard.h); a missing library raises ImportError.

``open_csr_shards`` reads a ``.scx`` file's 256-byte header and full catalog
(docs/format.md sections 2-3; scx-format/src/header.rs, catalog.rs) and returns
the CSR shard sections of X (section_type 4), ready for the decoder, so a
file goes from disk to a resident ``Matrix`` without a host CSR:

    csr = decode_csr_shards_to_device(open_csr_shards("cells.scx"))
    m = Matrix(csr.indptr, csr.indices, csr.data, csr.shape, target_sum=1e4, log1p=True)
"""
from __future__ import annotations

import ctypes
import mmap
import os
import struct
from dataclasses import dataclass, field

import numpy as np

from . import _lib

CODEC_NAMES = {0: "none", 1: "scx1", 2: "zstd", 3: "lz4shuffle", 4: "pcodec", 5: "shufdelta"}
SECTION_CSR_SHARD = 4

_FILE_HEADER = struct.Struct("<4sHHIQQQIIIBBBBQQQQQQQQQIQQ")  # 144 bytes used of the 256
_CATALOG_HEAD = struct.Struct("<HQQQI")


@dataclass
class GpuCsr:
    """Device CSR: indptr (n_rows + 1, int64), indices (nnz, int32), data (nnz, float32)."""
    indptr: object
    indices: object
    data: object
    shape: tuple
    stats: dict = field(default_factory=dict)

    @property
    def nnz(self) -> int:
        return int(self.indices.numel())


@dataclass
class Section:
    name: str
    offset: int
    length: int
    section_type: int
    modality_id: int


def _as_bytes(b) -> bytes | memoryview:
    return b if isinstance(b, (bytes, bytearray, memoryview)) else bytes(b)


def shard_info(shard) -> dict:
    """The shard header (and a framed shard's block index), checked."""
    b = np.frombuffer(_as_bytes(shard), np.uint8)
    info = _lib.ShardInfo()
    _lib.check(_lib.lib().scxshard_read_info(ctypes.c_void_p(b.ctypes.data), b.size, ctypes.byref(info)))
    out = {k: int(getattr(info, k)) for k, _ in _lib.ShardInfo._fields_}
    out["codec"] = CODEC_NAMES.get(out["codec_id"], str(out["codec_id"]))
    return out


def decode_csr_shards_to_device(shards, device: int = 0, stream=None) -> GpuCsr:
    """Decode shards (row-stacked in the given order) into one device CSR.

    All shards must declare the same column count; the result has shape
    (sum of shard rows, n_cols) (scx-gpu/src/gpu_csr_assemble.rs:70-90). A
    malformed shard raises ValueError with the reason."""
    import torch

    shards = [_as_bytes(s) for s in shards]
    if not shards:
        raise ValueError("decode_csr_shards_to_device: no CSR shards")
    infos = [shard_info(s) for s in shards]
    n_rows = sum(i["n_major"] for i in infos)
    nnz = sum(i["nnz"] for i in infos)
    n_cols = infos[0]["n_minor"]
    dev = torch.device("cuda", device)
    indptr = torch.empty(n_rows + 1, dtype=torch.int64, device=dev)
    indices = torch.empty(max(nnz, 1), dtype=torch.int32, device=dev)[:nnz]
    data = torch.empty(max(nnz, 1), dtype=torch.float32, device=dev)[:nnz]
    bufs = [np.frombuffer(s, np.uint8) for s in shards]
    ptrs = (ctypes.c_void_p * len(bufs))(*[b.ctypes.data for b in bufs])
    lens = (ctypes.c_uint64 * len(bufs))(*[b.size for b in bufs])
    st = _lib.ShardStats()
    torch.cuda.synchronize(dev)  # the outputs were allocated on torch's stream
    _lib.check(_lib.lib().scxshard_decode_csr(ctypes.cast(ptrs, ctypes.c_void_p), ctypes.cast(lens, ctypes.c_void_p),
                                              len(bufs), _lib.ptr(indptr), _lib.ptr(indices), _lib.ptr(data), n_rows,
                                              nnz, device, stream, ctypes.byref(st)))
    stats = {k: (float(getattr(st, k)) if t is ctypes.c_double else int(getattr(st, k)))
             for k, t in _lib.ShardStats._fields_}
    stats["fully_device_decoded"] = bool(stats["fully_device_decoded"])
    return GpuCsr(indptr, indices, data, (n_rows, n_cols), stats)


def decode_shard_gpu(shard, device: int = 0, stream=None) -> GpuCsr:
    """One shard onto the device (scx-gpu/src/shard_decode.rs:251-285)."""
    return decode_csr_shards_to_device([shard], device=device, stream=stream)


_VENC = {np.dtype(np.uint8): 0, np.dtype(np.uint16): 1, np.dtype(np.uint32): 2, np.dtype(np.float32): 3,
         np.dtype(np.float16): 4}


def encode_shard(indptr, indices, values, codec: str = "scx1", n_minor: int = 0, global_offset: int = 0,
                 row_group_rows: int = 256, index_dtype: int | None = None) -> bytes:
    """One CSR shard section (scxshard_encode): Scx1 (integer values) or None.
    ``row_group_rows`` > 0 frames it (shard format 2), as the reference
    writer does by default at 256 rows."""
    ip = np.ascontiguousarray(indptr, dtype=np.int64)
    ix = np.ascontiguousarray(indices, dtype=np.uint32)
    vals = np.ascontiguousarray(values)
    if vals.dtype not in _VENC:
        raise ValueError(f"values must be uint8/16/32, float32 or float16, not {vals.dtype}")
    if index_dtype is None:
        index_dtype = 0 if (n_minor and n_minor <= 65536) or (not n_minor and (ix.size == 0 or ix.max() <= 65535)) else 1
    cid = {"none": 0, "scx1": 1}[codec]
    args = [_lib.ptr(ip), len(ip) - 1, _lib.ptr(ix) if ix.size else None, _lib.ptr(vals) if vals.size else None,
            _VENC[vals.dtype], cid, index_dtype, n_minor, global_offset, row_group_rows]
    n = ctypes.c_uint64()
    _lib.check(_lib.lib().scxshard_encode(*args, None, 0, ctypes.byref(n)))
    out = np.empty(n.value, np.uint8)
    _lib.check(_lib.lib().scxshard_encode(*args, _lib.ptr(out), n.value, ctypes.byref(n)))
    return out.tobytes()


def encode_csr_shards(indptr, indices, values, n_cols: int, rows_per_shard: int = 16384, codec: str = "scx1",
                      row_group_rows: int = 256) -> list:
    """A CSR split into shards of ``rows_per_shard`` rows (the reference
    writer's DEFAULT_SHARD_TARGET_ROWS is 16,384), each framed at
    ``row_group_rows``."""
    ip = np.asarray(indptr, np.int64)
    n = len(ip) - 1
    out = []
    for a in range(0, max(n, 1), rows_per_shard):
        b = min(n, a + rows_per_shard)
        lo, hi = int(ip[a]), int(ip[b])
        out.append(encode_shard(ip[a:b + 1] - lo, indices[lo:hi], values[lo:hi], codec, n_cols, a,
                                row_group_rows if b > a else 0))
    return out


def read_catalog(path_or_bytes) -> tuple[dict, list[Section]]:
    """File header fields and the active full catalog's entries of a ``.scx``
    file (docs/format.md sections 2-3)."""
    if isinstance(path_or_bytes, (str, os.PathLike)):
        with open(path_or_bytes, "rb") as f:
            buf = mmap.mmap(f.fileno(), 0, access=mmap.ACCESS_READ)
    else:
        buf = _as_bytes(path_or_bytes)
    if len(buf) < 256:
        raise ValueError("not an SCX file: shorter than the 256-byte header")
    h = _FILE_HEADER.unpack_from(buf, 0)
    names = ("magic", "format_version", "header_length", "flags", "n_obs", "n_vars", "nnz", "n_csr_shards",
             "n_csc_shards", "shard_target_rows", "codec_id", "index_dtype", "endian", "reserved_padding",
             "root_catalog_offset", "root_catalog_length", "full_catalog_offset", "full_catalog_length",
             "manifest_sequence", "prev_catalog_offset", "file_checksum", "front_catalog_offset",
             "front_catalog_length", "n_modalities", "modality_table_offset", "modality_table_length")
    hdr = dict(zip(names, h))
    if hdr["magic"] != b"SCX\x01":
        raise ValueError("not an SCX file: bad magic")
    if hdr["endian"] != 0:
        raise ValueError("big-endian SCX files are not supported")
    off, ln = hdr["full_catalog_offset"], hdr["full_catalog_length"]
    if off + ln > len(buf) or ln < _CATALOG_HEAD.size:
        raise ValueError("full catalog lies outside the file")
    version, _seq, _prev, _nobs, n_entries = _CATALOG_HEAD.unpack_from(buf, off)
    p, end = off + _CATALOG_HEAD.size, off + ln
    sections = []
    for _ in range(n_entries):
        (nl,) = struct.unpack_from("<H", buf, p)
        p += 2
        name = bytes(buf[p:p + nl]).decode("utf-8")
        p += nl
        so, sl, st = struct.unpack_from("<QQB", buf, p)
        p += 17 + 32  # + BLAKE3 checksum
        mid = 0
        if version >= 2:
            mid = buf[p]
            p += 1
        (stl,) = struct.unpack_from("<H", buf, p)
        p += 2 + stl
        if p > end:
            raise ValueError("full catalog entry runs past the catalog")
        sections.append(Section(name, so, sl, st, mid))
    return hdr, sections


def open_csr_shards(path_or_bytes, modality_id: int | None = None) -> list:
    """The X CSR shard sections of a ``.scx`` file, ordered by their first
    row. ``modality_id`` picks one modality of a multimodal file (0 = global)."""
    if isinstance(path_or_bytes, (str, os.PathLike)):
        with open(path_or_bytes, "rb") as f:
            buf = f.read()
    else:
        buf = _as_bytes(path_or_bytes)
    _, sections = read_catalog(buf)
    out = []
    for s in sections:
        if s.section_type != SECTION_CSR_SHARD:
            continue
        if modality_id is not None and s.modality_id != modality_id:
            continue
        if s.offset + s.length > len(buf):
            raise ValueError(f"section {s.name} lies outside the file")
        sec = memoryview(buf)[s.offset:s.offset + s.length]
        (start,) = struct.unpack_from("<Q", sec, 28) if s.length >= 36 else (0,)
        out.append((start, sec))
    out.sort(key=lambda t: t[0])
    return [sec for _, sec in out]
