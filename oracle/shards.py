"""ctypes wrapper over the shard-codec half of ``liborc.so`` — test
infrastructure only (tests/, smoke(), bench.py's cpu_baseline leg).

``scxshard_oracle.c`` restates scx-codec's scalar decoders/encoders (Rice,
Delta-Golomb, FOR-BP with the BitPacker4x layout, byte shuffle/delta) and
scx-format's 76-byte shard header + block index; zstd and LZ4 frames go
through the system libzstd.so.1 / liblz4.so.1.

Parity status:
  * Rice, Delta-Golomb, FOR-BP scalar rows, shard header layout — pinned by
    the byte-exact vectors of scx-codec/tests/reference_vectors.rs and by the
    reference's golden ``.scx`` files (tests/reference_files/golden_*), whose
    JSON twins hold the decoded CSR.
  * BitPacker4x (rows of >= 128 entries) — no reference fixture holds such a
    row; pinned by the layout the reference's device decoder states
    (scx-gpu/kernels/forbp_decode_bp4x.cu:8-16), by its reduction to the plain
    little-endian u32 array at 32 bits, and by round trips.
  * byte shuffle / delta — pinned by scx-codec/src/byte_delta.rs's worked
    example and the golden lz4shuffle / zstd files.
"""
from __future__ import annotations

import ctypes

import numpy as np

from . import lib as _orc_lib

_i64 = ctypes.c_int64
_i32 = ctypes.c_int32
_p = ctypes.c_void_p

CODECS = {"none": 0, "scx1": 1, "zstd": 2, "lz4shuffle": 3, "pcodec": 4, "shufdelta": 5}
VALUE_ENCODINGS = {"u8": 0, "u16": 1, "u32": 2, "f32": 3, "f16": 4}
_VDTYPE = {0: np.uint8, 1: np.uint16, 2: np.uint32, 3: np.float32, 4: np.float16}


class ShardOracleError(ValueError):
    pass


def _lib():
    lib = _orc_lib()
    if not getattr(lib, "_sho_ready", False):
        for name in ("sho_rice_encode", "sho_rice_decode", "sho_dg_encode", "sho_dg_decode", "sho_forbp_encode",
                     "sho_forbp_decode", "sho_encode_shard"):
            getattr(lib, name).restype = _i64
        for name in ("sho_shard_info", "sho_decode_shard", "sho_have_zstd", "sho_have_lz4"):
            getattr(lib, name).restype = _i32
        lib.sho_last_error.restype = ctypes.c_char_p
        lib._sho_ready = True
    return lib


def _ptr(a):
    return _p(a.ctypes.data)


def _err(what):
    raise ShardOracleError(f"{what}: {_lib().sho_last_error().decode()}")


def have_zstd() -> bool:
    return bool(_lib().sho_have_zstd())


def have_lz4() -> bool:
    return bool(_lib().sho_have_lz4())


def rice_encode(values) -> bytes:
    v = np.ascontiguousarray(values, dtype=np.uint32)
    n = _lib().sho_rice_encode(_ptr(v), len(v), None, 0)
    if n < 0:
        _err("rice_encode")
    out = np.zeros(max(n, 1), np.uint8)
    n = _lib().sho_rice_encode(_ptr(v), len(v), _ptr(out), len(out))
    return out[:n].tobytes()


def rice_decode(data: bytes, n: int) -> np.ndarray:
    buf = np.frombuffer(data, np.uint8).copy() if data else np.zeros(1, np.uint8)
    out = np.zeros(max(n, 1), np.uint32)
    if _lib().sho_rice_decode(_ptr(buf), len(data), n, _ptr(out)) < 0:
        _err("rice_decode")
    return out[:n]


def delta_golomb_encode(indptr) -> bytes:
    a = np.ascontiguousarray(indptr, dtype=np.uint64)
    n = _lib().sho_dg_encode(_ptr(a), len(a), None, 0)
    if n < 0:
        _err("delta_golomb_encode")
    out = np.zeros(max(n, 1), np.uint8)
    n = _lib().sho_dg_encode(_ptr(a), len(a), _ptr(out), len(out))
    return out[:n].tobytes()


def delta_golomb_decode(data: bytes, n: int) -> np.ndarray:
    buf = np.frombuffer(data, np.uint8).copy() if data else np.zeros(1, np.uint8)
    out = np.zeros(max(n, 1), np.uint64)
    if _lib().sho_dg_decode(_ptr(buf), len(data), n, _ptr(out)) < 0:
        _err("delta_golomb_decode")
    return out[:n]


def forbp_encode(indices, row_lengths, u16: bool) -> bytes:
    ix = np.ascontiguousarray(indices, dtype=np.uint32)
    rl = np.ascontiguousarray(row_lengths, dtype=np.uint64)
    ixp = _ptr(ix) if len(ix) else _ptr(np.zeros(1, np.uint32))
    n = _lib().sho_forbp_encode(ixp, _ptr(rl), len(rl), int(u16), None, 0)
    if n < 0:
        _err("forbp_encode")
    out = np.zeros(max(n, 1), np.uint8)
    n = _lib().sho_forbp_encode(ixp, _ptr(rl), len(rl), int(u16), _ptr(out), len(out))
    return out[:n].tobytes()


def forbp_decode(data: bytes, n_rows: int, u16: bool, nnz_hint: int = 0, cap: int | None = None):
    buf = np.frombuffer(data, np.uint8).copy() if data else np.zeros(1, np.uint8)
    cap = cap if cap is not None else (nnz_hint if nnz_hint else 1 << 24)
    idx = np.zeros(max(cap, 1), np.uint32)
    rl = np.zeros(max(n_rows, 1), np.uint64)
    n = _lib().sho_forbp_decode(_ptr(buf), len(data), n_rows, nnz_hint, int(u16), _ptr(idx), cap, _ptr(rl))
    if n < 0:
        _err("forbp_decode")
    return idx[:n], rl[:n_rows]


def bp4x_pack(values128, bits: int) -> bytes:
    v = np.ascontiguousarray(values128, dtype=np.uint32)
    assert v.shape == (128,)
    out = np.zeros(bits * 16 if bits else 1, np.uint8)
    _lib().sho_bp4x_pack(_ptr(v), bits, _ptr(out))
    return out[: bits * 16].tobytes()


def bp4x_unpack(data: bytes, bits: int) -> np.ndarray:
    buf = np.frombuffer(data, np.uint8).copy()
    out = np.zeros(128, np.uint32)
    _lib().sho_bp4x_unpack(_ptr(buf), bits, _ptr(out))
    return out


def byte_shuffle(raw: bytes, width: int) -> bytes:
    a = np.frombuffer(raw, np.uint8).copy()
    out = np.zeros_like(a)
    _lib().sho_byte_shuffle(_ptr(a), len(a) // width, width, _ptr(out))
    return out.tobytes()


def byte_unshuffle(raw: bytes, width: int) -> bytes:
    a = np.frombuffer(raw, np.uint8).copy()
    out = np.zeros_like(a)
    _lib().sho_byte_unshuffle(_ptr(a), len(a) // width, width, _ptr(out))
    return out.tobytes()


def byte_delta(raw: bytes, planes: int, n: int) -> bytes:
    a = np.frombuffer(raw, np.uint8).copy()
    _lib().sho_byte_delta(_ptr(a), planes, _i64(n))
    return a.tobytes()


def byte_undelta(raw: bytes, planes: int, n: int) -> bytes:
    a = np.frombuffer(raw, np.uint8).copy()
    _lib().sho_byte_undelta(_ptr(a), planes, _i64(n))
    return a.tobytes()


def encode_shard(indptr, indices, values, codec: str | int = "scx1", value_encoding: str | int | None = None,
                 u16: bool | None = None, n_minor: int = 0, global_offset: int = 0, row_group_rows: int = 0) -> bytes:
    """One CSR shard section (header, sub-streams, block index); framed (shard
    v2) when ``row_group_rows`` > 0."""
    codec = CODECS[codec] if isinstance(codec, str) else int(codec)
    ip = np.ascontiguousarray(indptr, dtype=np.uint64)
    ix = np.ascontiguousarray(indices, dtype=np.uint32)
    vals = np.asarray(values)
    if value_encoding is None:
        value_encoding = {np.dtype(np.uint8): 0, np.dtype(np.uint16): 1, np.dtype(np.uint32): 2,
                          np.dtype(np.float32): 3, np.dtype(np.float16): 4}[vals.dtype]
    venc = VALUE_ENCODINGS[value_encoding] if isinstance(value_encoding, str) else int(value_encoding)
    vals = np.ascontiguousarray(vals, dtype=_VDTYPE[venc])
    if u16 is None:
        u16 = n_minor <= 65536 and (len(ix) == 0 or int(ix.max()) <= 65535)
    raw = vals.view(np.uint8) if len(vals) else np.zeros(1, np.uint8)
    ixp = ix if len(ix) else np.zeros(1, np.uint32)
    args = (_ptr(ip), len(ip) - 1, _ptr(ixp), _ptr(raw), codec, venc, int(u16), n_minor, ctypes.c_uint64(global_offset),
            row_group_rows)
    n = _lib().sho_encode_shard(*args, None, 0)
    if n < 0:
        _err("encode_shard")
    out = np.zeros(n, np.uint8)
    if _lib().sho_encode_shard(*args, _ptr(out), n) != n:
        _err("encode_shard")
    return out.tobytes()


def shard_info(shard: bytes) -> dict:
    buf = np.frombuffer(shard, np.uint8)
    vals = [_i64(), _i64(), _i64(), _i32(), _i32(), _i32(), _i32()]
    if _lib().sho_shard_info(_ptr(buf), len(buf), *[ctypes.byref(v) for v in vals]) != 0:
        _err("shard_info")
    keys = ("n_major", "n_minor", "nnz", "codec", "value_encoding", "index_dtype", "shard_format_version")
    return dict(zip(keys, (v.value for v in vals)))


def decode_shard(shard: bytes):
    """decode_shard_scipy: (i64 indptr, i32 indices, f32 data)."""
    buf = np.frombuffer(shard, np.uint8)
    info = shard_info(shard)
    n, nnz = info["n_major"], info["nnz"]
    if nnz > 8 * len(shard) + (1 << 20):
        raise ShardOracleError(f"decode_shard: header nnz {nnz} is implausible for {len(shard)} bytes")
    indptr = np.zeros(n + 1, np.int64)
    idx = np.zeros(max(nnz, 1), np.int32)
    data = np.zeros(max(nnz, 1), np.float32)
    if _lib().sho_decode_shard(_ptr(buf), len(buf), _ptr(indptr), _ptr(idx), _ptr(data)) != 0:
        _err("decode_shard")
    return indptr, idx[:nnz], data[:nnz]
