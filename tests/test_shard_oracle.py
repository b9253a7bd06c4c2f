"""The shard-codec oracle (oracle/scxshard_oracle.c) against the reference's
own known answers, and the host-side .scx catalog reader. No GPU.

Pins:
* byte-exact vectors of scx-codec/tests/reference_vectors.rs (Rice :16-68,
  FOR-BP :73-121, Delta-Golomb :127-170) and byte_delta.rs's worked example;
* the reference's golden .scx files (tests/reference_files/golden_*, copied to
  tests/golden/scx by tests/golden/make_scx_fixtures.py): every codec the
  oracle implements decodes them to the CSR their JSON twins hold, and the
  oracle's encoder rebuilds the Scx1 / None shards byte for byte (all but the
  8-byte checksum);
* BitPacker4x: at 32 bits the layout is the plain little-endian u32 array; at
  1 bit value p lands on bit (p >> 2) of lane p & 3's first word; round trips.
"""
from __future__ import annotations

import os
import struct

import numpy as np
import pytest

from oracle import shards as S
from scx_amd import shards as host

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "scx")
GOLDEN = sorted(f[:-4] for f in os.listdir(GOLD) if f.startswith("golden_") and f.endswith(".scx"))
_VD = {"u8": np.uint8, "u16": np.uint16, "u32": np.uint32, "f32": np.float32}


def _golden(name):
    buf = open(os.path.join(GOLD, name + ".scx"), "rb").read()
    exp = np.load(os.path.join(GOLD, name + ".npz"))
    return buf, exp


# ---- reference_vectors.rs ----
def test_rice_reference_vectors():
    assert S.rice_encode([1, 1, 1, 1]) == bytes([0x00, 0x00])
    assert S.rice_encode([1, 2]) == bytes([0x00, 0x02])
    assert S.rice_encode([4, 5, 6, 7]) == bytes([0x01, 0x9D, 0x3D])
    np.testing.assert_array_equal(S.rice_decode(bytes([0x01, 0x9D, 0x3D]), 4), [4, 5, 6, 7])


def test_forbp_reference_vectors():
    e = S.forbp_encode([0, 5, 10], [3], True)
    assert e[0:4] == bytes([3, 0, 0, 0]) and e[4:6] == bytes([1, 0]) and e[6] == 3
    assert e[7:9] == bytes([0, 0]) and e[9] == 3
    idx, rl = S.forbp_decode(e, 1, True)
    assert idx.tolist() == [0, 5, 10] and rl.tolist() == [3]
    e = S.forbp_encode([42], [0, 1], True)
    assert e == bytes([1, 0, 0, 0, 2, 0, 0, 1, 42, 0, 0])
    idx, rl = S.forbp_decode(e, 2, True)
    assert idx.tolist() == [42] and rl.tolist() == [0, 1]


def test_delta_golomb_reference_vectors():
    e = S.delta_golomb_encode([0, 150, 280, 500])
    assert e[:8] == bytes(8) and e[8] == 6
    assert S.delta_golomb_decode(e, 4).tolist() == [0, 150, 280, 500]
    assert S.delta_golomb_encode([42]) == bytes([42, 0, 0, 0, 0, 0, 0, 0])
    e = S.delta_golomb_encode([0, 0, 0])
    assert e[8] == 0 and e[9] == 0
    assert S.delta_golomb_decode(e, 3).tolist() == [0, 0, 0]


def test_byte_delta_worked_example():
    orig = bytes([10, 12, 12, 20, 0, 255, 1, 3])
    d = S.byte_delta(orig, 2, 4)
    assert d == bytes([10, 2, 0, 8, 0, 255, 2, 2])
    assert S.byte_undelta(d, 2, 4) == orig


def test_rice_k_selection_and_block_boundaries():
    # median of [0..7] (floor) = 3 -> k = 1 (rice.rs:185-195); 257 values = two blocks
    assert S.rice_encode(list(range(1, 9)))[0] == 1
    vals = np.arange(1, 258, dtype=np.uint32)
    e = S.rice_encode(vals)
    np.testing.assert_array_equal(S.rice_decode(e, 257), vals)
    with pytest.raises(S.ShardOracleError, match="reserved high nibble"):
        S.rice_decode(bytes([0x10, 0x00]), 1)
    with pytest.raises(S.ShardOracleError, match="truncated"):
        S.rice_decode(bytes([0x00]), 1)


# ---- BitPacker4x ----
def test_bp4x_32_bits_is_the_little_endian_array():
    v = np.random.default_rng(0).integers(0, 2**32, 128, dtype=np.uint64).astype(np.uint32)
    assert S.bp4x_pack(v, 32) == v.astype("<u4").tobytes()


def test_bp4x_1_bit_lane_layout():
    for p in (0, 1, 2, 3, 4, 5, 31, 64, 127):
        v = np.zeros(128, np.uint32)
        v[p] = 1
        b = S.bp4x_pack(v, 1)
        lane, slot = p & 3, p >> 2
        expect = bytearray(16)
        expect[4 * lane + slot // 8] = 1 << (slot % 8)
        assert b == bytes(expect), p


@pytest.mark.parametrize("bits", list(range(1, 33)))
def test_bp4x_round_trip(bits):
    rng = np.random.default_rng(bits)
    v = rng.integers(0, 2**bits, 128, dtype=np.uint64).astype(np.uint32)
    np.testing.assert_array_equal(S.bp4x_unpack(S.bp4x_pack(v, bits), bits), v)


# ---- golden .scx files ----
@pytest.mark.parametrize("name", GOLDEN)
def test_golden_file_decodes(name):
    buf, exp = _golden(name)
    shards = host.open_csr_shards(buf)
    assert len(shards) == 1
    if "pcodec" in name:
        with pytest.raises(S.ShardOracleError, match="pcodec"):
            S.decode_shard(bytes(shards[0]))
        return
    ip, ix, d = S.decode_shard(bytes(shards[0]))
    np.testing.assert_array_equal(ip, exp["indptr"])
    np.testing.assert_array_equal(ix, exp["indices"])
    np.testing.assert_array_equal(d, exp["data"])
    info = S.shard_info(bytes(shards[0]))
    assert (info["n_major"], info["n_minor"], info["nnz"]) == (100, 50, 1506)


@pytest.mark.parametrize("name", [n for n in GOLDEN if n.split("_")[1] in ("scx1", "none")])
def test_encoder_rebuilds_golden_shards(name):
    buf, exp = _golden(name)
    sh = bytes(host.open_csr_shards(buf)[0])
    venc = name.split("_")[2]
    enc = S.encode_shard(exp["indptr"], exp["indices"], exp["data"].astype(_VD[venc]), name.split("_")[1], venc,
                         u16=True, n_minor=50)
    assert len(enc) == len(sh)
    assert enc[:68] == sh[:68] and enc[76:] == sh[76:]  # bytes 68..76: BLAKE3 checksum


def test_multimodal_catalog_v2():
    buf = open(os.path.join(GOLD, "v2_multimodal_citeseq.scx"), "rb").read()
    exp = np.load(os.path.join(GOLD, "v2_multimodal_citeseq.npz"))
    hdr, secs = host.read_catalog(buf)
    assert hdr["n_modalities"] == 2
    for mid in (1, 2):
        shards = host.open_csr_shards(buf, modality_id=mid)
        assert len(shards) == 1
        ip, ix, d = S.decode_shard(bytes(shards[0]))
        np.testing.assert_array_equal(ip, exp[f"m{mid}_indptr"])
        np.testing.assert_array_equal(ix, exp[f"m{mid}_indices"])
        np.testing.assert_array_equal(d, exp[f"m{mid}_data"])


# ---- synthetic round trips (the reference's GPU test shapes) ----
def dense_rows(row_nnzs, n_cols, seed=7):
    """Sorted distinct columns per row and UMI-like values (1..5, some large)."""
    rng = np.random.default_rng(seed)
    ip, ix, vv = [0], [], []
    for nz in row_nnzs:
        cols = np.sort(rng.choice(n_cols, size=nz, replace=False)).astype(np.uint32)
        ix.append(cols)
        v = rng.integers(1, 6, nz)
        large = rng.random(nz) < 0.05
        v[large] = rng.integers(6, 60000, int(large.sum()))
        vv.append(v)
        ip.append(ip[-1] + nz)
    ix = np.concatenate(ix) if ix else np.zeros(0, np.uint32)
    vv = np.concatenate(vv).astype(np.uint16) if vv else np.zeros(0, np.uint16)
    return np.asarray(ip, np.uint64), ix, vv


ROWS = [0, 1, 5, 130, 256, 7, 0, 384, 200, 3, 128, 129, 512, 50]


@pytest.mark.parametrize("codec", ["scx1", "none", "zstd", "lz4shuffle", "shufdelta"])
@pytest.mark.parametrize("n_cols", [4000, 70000])
@pytest.mark.parametrize("framing", [0, 4])
def test_round_trip(codec, n_cols, framing):
    ip, ix, vv = dense_rows(ROWS, n_cols)
    sh = S.encode_shard(ip, ix, vv, codec, "u16", u16=n_cols <= 65536, n_minor=n_cols, row_group_rows=framing)
    info = S.shard_info(sh)
    assert info["shard_format_version"] == (2 if framing else 1)
    gi, gx, gd = S.decode_shard(sh)
    np.testing.assert_array_equal(gi, ip.astype(np.int64))
    np.testing.assert_array_equal(gx, ix.astype(np.int32))
    np.testing.assert_array_equal(gd, vv.astype(np.float32))


def test_float_values_shufdelta_and_f16():
    ip, ix, vv = dense_rows([3, 10, 130, 0, 256, 5], 4000)
    f32 = vv.astype(np.float32) + 0.5
    for codec in ("shufdelta", "zstd", "none", "lz4shuffle"):
        sh = S.encode_shard(ip, ix, f32, codec, "f32", n_minor=4000, row_group_rows=4)
        np.testing.assert_array_equal(S.decode_shard(sh)[2], f32)
    f16 = (vv.astype(np.float32) / 7).astype(np.float16)
    sh = S.encode_shard(ip, ix, f16, "none", "f16", n_minor=4000)
    np.testing.assert_array_equal(S.decode_shard(sh)[2], f16.astype(np.float32))
    with pytest.raises(S.ShardOracleError, match="float"):
        S.encode_shard(ip, ix, f32, "scx1", "f32", n_minor=4000)


def test_malformed_shards_are_rejected():
    ip, ix, vv = dense_rows([3, 10, 130, 2], 4000)
    sh = bytearray(S.encode_shard(ip, ix, vv, "scx1", "u16", n_minor=4000))
    bad = bytearray(sh)
    bad[0] = ord("X")
    with pytest.raises(S.ShardOracleError, match="magic"):
        S.decode_shard(bytes(bad))
    # an index bound below the largest column
    small = S.encode_shard(ip, ix, vv, "scx1", "u16", n_minor=int(ix.max()))
    with pytest.raises(S.ShardOracleError, match="out of range"):
        S.decode_shard(small)
    # Rice block header with the reserved nibble set
    vo = struct.unpack_from("<I", sh, 52)[0]
    bad = bytearray(sh)
    bad[vo] |= 0x30
    with pytest.raises(S.ShardOracleError, match="reserved high nibble"):
        S.decode_shard(bytes(bad))
    # framed block index whose nnz disagrees with the header
    fr = bytearray(S.encode_shard(ip, ix, vv, "scx1", "u16", n_minor=4000, row_group_rows=2))
    bo = struct.unpack_from("<I", fr, 60)[0]
    struct.pack_into("<I", fr, bo + 4 + 18, struct.unpack_from("<I", fr, bo + 4 + 18)[0] + 1)
    with pytest.raises(S.ShardOracleError):
        S.decode_shard(bytes(fr))


# ---- the library's shard writer (host code of libscxpca.so) ----
@pytest.mark.parametrize("codec", ["scx1", "none"])
@pytest.mark.parametrize("n_cols", [4000, 70000])
@pytest.mark.parametrize("framing", [0, 4, 256])
@pytest.mark.parametrize("venc", ["u8", "u16", "u32"])
def test_library_writer_matches_oracle_encoder(codec, n_cols, framing, venc):
    rng = np.random.default_rng(5)
    rows = [int(x) for x in rng.integers(0, 700, 30)] + ROWS
    ip, ix, vv = dense_rows(rows, n_cols, seed=9)
    vv = ((vv % 250) + 1).astype(_VD[venc]) if venc == "u8" else vv.astype(_VD[venc])
    u16 = n_cols <= 65536
    lib = host.encode_shard(ip.astype(np.int64), ix, vv, codec, n_minor=n_cols, global_offset=12345,
                            row_group_rows=framing, index_dtype=0 if u16 else 1)
    ref = S.encode_shard(ip, ix, vv, codec, venc, u16=u16, n_minor=n_cols, global_offset=12345,
                         row_group_rows=framing)
    assert lib == ref


def test_library_writer_golden_and_errors():
    buf, exp = _golden("golden_scx1_u16")
    sh = bytes(host.open_csr_shards(buf)[0])
    lib = host.encode_shard(exp["indptr"], exp["indices"], exp["data"].astype(np.uint16), "scx1", n_minor=50,
                            row_group_rows=0)
    assert lib[:68] == sh[:68] and lib[76:] == sh[76:]
    with pytest.raises(ValueError, match="float"):
        host.encode_shard([0, 1], [0], np.ones(1, np.float32), "scx1", n_minor=4)
    with pytest.raises(ValueError, match="non-zero"):
        host.encode_shard([0, 1], [0], np.zeros(1, np.uint16), "scx1", n_minor=4)
    with pytest.raises(ValueError, match="sorted"):
        host.encode_shard([0, 2], [3, 1], np.ones(2, np.uint16), "scx1", n_minor=4)
