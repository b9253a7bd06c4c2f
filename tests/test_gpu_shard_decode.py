"""On-GPU shard decode (scxshard_decode_csr) against the oracle and the
reference's golden files: bit-exact indptr / indices / data for every codec the
library decodes, framed and unframed shards, u16 and u32 indices, multi-shard
assembly, and the error paths. A 250,000-cell slice of the 1,000,000 x 2,000
workload goes through Scx1 shards of 16,384 rows framed at 256 rows (the
reference writer's defaults) and must decode to the source CSR and give the
same PCA as the source CSR."""
from __future__ import annotations

import os
import struct
import sys
from concurrent.futures import ThreadPoolExecutor

import numpy as np
import pytest

from oracle import shards as S
from scx_amd import shards as host
from scx_amd import synth
from tests.test_shard_oracle import GOLD, GOLDEN, ROWS, dense_rows

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _host(csr):
    return (csr.indptr.cpu().numpy(), csr.indices.cpu().numpy(), csr.data.cpu().numpy())


def _check(csr, ip, ix, d):
    gi, gx, gd = _host(csr)
    np.testing.assert_array_equal(gi, np.asarray(ip, np.int64))
    np.testing.assert_array_equal(gx, np.asarray(ix).astype(np.int32))
    np.testing.assert_array_equal(gd.view(np.uint32), np.asarray(d, np.float32).view(np.uint32))


@pytest.mark.parametrize("name", GOLDEN)
def test_golden_files(gpu, name):
    buf = open(os.path.join(GOLD, name + ".scx"), "rb").read()
    exp = np.load(os.path.join(GOLD, name + ".npz"))
    shards = host.open_csr_shards(buf)
    say(f"case: golden {name}")
    if "pcodec" in name:
        with pytest.raises(ValueError, match="pcodec"):
            host.decode_csr_shards_to_device(shards)
        return
    csr = host.decode_csr_shards_to_device(shards)
    assert csr.shape == (100, 50)
    _check(csr, exp["indptr"], exp["indices"], exp["data"])


def test_multimodal_file(gpu):
    buf = open(os.path.join(GOLD, "v2_multimodal_citeseq.scx"), "rb").read()
    exp = np.load(os.path.join(GOLD, "v2_multimodal_citeseq.npz"))
    for mid in (1, 2):
        say(f"case: multimodal modality {mid}")
        csr = host.decode_csr_shards_to_device(host.open_csr_shards(buf, modality_id=mid))
        _check(csr, exp[f"m{mid}_indptr"], exp[f"m{mid}_indices"], exp[f"m{mid}_data"])


_VALS = {"u8": np.uint8, "u16": np.uint16, "u32": np.uint32}


@pytest.mark.parametrize("codec", ["scx1", "none", "zstd", "lz4shuffle", "shufdelta"])
@pytest.mark.parametrize("n_cols", [4000, 70000])
@pytest.mark.parametrize("framing", [0, 4])
@pytest.mark.parametrize("venc", ["u8", "u16", "u32"])
def test_codecs_match_oracle(gpu, codec, n_cols, framing, venc):
    ip, ix, vv = dense_rows(ROWS, n_cols)
    vv = (vv % 250 + 1) if venc == "u8" else vv.astype(np.uint32) * (7 if venc == "u32" else 1)
    vv = vv.astype(_VALS[venc])
    sh = S.encode_shard(ip, ix, vv, codec, venc, u16=n_cols <= 65536, n_minor=n_cols, row_group_rows=framing)
    say(f"case: {codec} n_cols={n_cols} framing={framing} {venc} ({len(sh)} bytes)")
    ref = S.decode_shard(sh)
    csr = host.decode_shard_gpu(sh)
    _check(csr, *ref)
    _check(csr, ip.astype(np.int64), ix, vv.astype(np.float32))
    st = csr.stats
    assert st["n_groups"] == (4 if framing else 1)
    assert st["fully_device_decoded"] == (codec in ("scx1", "none"))


def test_float_values(gpu):
    ip, ix, vv = dense_rows([3, 10, 130, 0, 256, 5], 4000)
    f32 = vv.astype(np.float32) + 0.5
    for codec in ("shufdelta", "zstd", "none", "lz4shuffle"):
        say(f"case: f32 values {codec}")
        sh = S.encode_shard(ip, ix, f32, codec, "f32", n_minor=4000, row_group_rows=4)
        _check(host.decode_shard_gpu(sh), *S.decode_shard(sh))
    f16 = np.concatenate([(vv.astype(np.float32) / 7).astype(np.float16)[:-3],
                          np.array([6.0e-8, 65504.0, -2.5], np.float16)])
    say("case: f16 values none")
    sh = S.encode_shard(ip, ix, f16, "none", "f16", n_minor=4000)
    _check(host.decode_shard_gpu(sh), ip.astype(np.int64), ix, f16.astype(np.float32))


def test_large_rows_and_wide_frames(gpu):
    """Rows up to 5,000 entries (BitPacker4x chunks plus tails, frame widths up
    to 17 bits) and a delta-coded plane spanning many 2,048-element chunks."""
    rng = np.random.default_rng(3)
    rows = [int(x) for x in rng.integers(0, 5000, 40)] + [4999, 128, 127, 1, 0]
    ip, ix, vv = dense_rows(rows, 70000, seed=11)
    for codec, g in (("scx1", 0), ("scx1", 16), ("shufdelta", 0), ("shufdelta", 64), ("lz4shuffle", 0)):
        say(f"case: large rows {codec} framing={g}")
        sh = S.encode_shard(ip, ix, vv, codec, "u16", u16=False, n_minor=70000, row_group_rows=g)
        _check(host.decode_shard_gpu(sh), ip.astype(np.int64), ix, vv.astype(np.float32))


def test_multi_shard_assembly(gpu):
    ip, ix, vv = dense_rows(ROWS * 6, 4000, seed=5)
    n = len(ip) - 1
    cuts = [0, 10, 11, 40, 60, 84]
    codecs = [("scx1", 4), ("shufdelta", 0), ("none", 0), ("zstd", 3), ("lz4shuffle", 0)]
    shards = []
    for (a, b), (codec, g) in zip(zip(cuts[:-1], cuts[1:]), codecs):
        lip = ip[a:b + 1] - ip[a]
        sl = slice(int(ip[a]), int(ip[b]))
        shards.append(S.encode_shard(lip, ix[sl], vv[sl], codec, "u16", n_minor=4000, global_offset=a,
                                     row_group_rows=g))
    assert cuts[-1] == n
    say("case: five shards, mixed codecs")
    csr = host.decode_csr_shards_to_device(shards)
    assert csr.shape == (n, 4000)
    _check(csr, ip.astype(np.int64), ix, vv.astype(np.float32))
    assert csr.stats["n_shards_scx1_gpu"] == 1 and csr.stats["n_shards_shufdelta_gpu"] == 1
    assert csr.stats["n_shards_other_gpu"] == 3 and not csr.stats["fully_device_decoded"]


def test_errors(gpu):
    ip, ix, vv = dense_rows([3, 10, 130, 2], 4000)
    good = S.encode_shard(ip, ix, vv, "scx1", "u16", n_minor=4000)
    say("case: error paths")
    other = S.encode_shard(ip, ix, vv, "scx1", "u16", n_minor=5000)
    with pytest.raises(ValueError, match="column count"):
        host.decode_csr_shards_to_device([good, other])
    small = S.encode_shard(ip, ix, vv, "scx1", "u16", n_minor=int(ix.max()))
    with pytest.raises(ValueError, match="out of range for n_minor"):
        host.decode_shard_gpu(small)
    vo = struct.unpack_from("<I", good, 52)[0]
    bad = bytearray(good)
    bad[vo] |= 0x30
    with pytest.raises(ValueError, match="reserved high nibble"):
        host.decode_shard_gpu(bytes(bad))
    bad = bytearray(good)
    bad[0] = ord("X")
    with pytest.raises(ValueError, match="magic"):
        host.decode_shard_gpu(bytes(bad))
    fr = bytearray(S.encode_shard(ip, ix, vv, "scx1", "u16", n_minor=4000, row_group_rows=2))
    bo = struct.unpack_from("<I", fr, 60)[0]
    struct.pack_into("<I", fr, bo + 4 + 18, struct.unpack_from("<I", fr, bo + 4 + 18)[0] + 1)
    with pytest.raises(ValueError, match="block index"):
        host.decode_shard_gpu(bytes(fr))
    zs = bytearray(S.encode_shard(ip, ix, vv, "zstd", "u16", n_minor=4000))
    xo, xl = struct.unpack_from("<II", zs, 44)
    zs[xo] ^= 0xFF  # the zstd frame magic
    with pytest.raises(ValueError, match="zstd frame"):
        host.decode_shard_gpu(bytes(zs))
    # the library still works after every refusal
    _check(host.decode_shard_gpu(good), ip.astype(np.int64), ix, vv.astype(np.float32))


def _encode_shards(x, rows_per_shard=16384, codec="scx1", row_group_rows=256):
    ip = x.indptr.astype(np.uint64)
    vals = x.data.astype(np.uint16)
    assert np.array_equal(vals.astype(np.float32), x.data)
    cuts = list(range(0, x.shape[0], rows_per_shard)) + [x.shape[0]]

    def one(ab):
        a, b = ab
        sl = slice(int(ip[a]), int(ip[b]))
        return S.encode_shard(ip[a:b + 1] - ip[a], x.indices[sl], vals[sl], codec, "u16", u16=True,
                              n_minor=x.shape[1], global_offset=a, row_group_rows=row_group_rows)

    with ThreadPoolExecutor(16) as ex:
        return list(ex.map(one, zip(cuts[:-1], cuts[1:])))


def test_workload_slice_scx1_round_trip_and_pca(gpu):
    x = synth.generate(synth.CFG2, n_obs=250_000)
    shards = _encode_shards(x)
    enc = sum(len(s) for s in shards)
    say(f"case: {len(shards)} Scx1 shards of 16384 rows (G=256), {x.nnz} nnz, {enc / 2**20:.1f} MiB encoded")
    csr = host.decode_csr_shards_to_device(shards)
    _check(csr, x.indptr, x.indices, x.data)
    st = csr.stats
    say(f"decode: host {st['host_seconds'] * 1e3:.1f} ms, upload {st['upload_seconds'] * 1e3:.1f} ms, "
        f"device {st['device_seconds'] * 1e3:.2f} ms, total {st['total_seconds'] * 1e3:.1f} ms")
    assert st["fully_device_decoded"] and st["n_shards_scx1_gpu"] == len(shards)
    kw = dict(n_comps=50, n_oversamples=10, n_power_iterations=3, seed=0, qr_method=1)
    m1 = gpu.Matrix(csr.indptr, csr.indices, csr.data, csr.shape, target_sum=1e4, log1p=True)
    m2 = gpu.Matrix(x.indptr, x.indices, x.data, x.shape, target_sum=1e4, log1p=True)
    try:
        a, b = m1.pca(**kw), m2.pca(**kw)
    finally:
        m1.close()
        m2.close()
    np.testing.assert_array_equal(a["sigma"], b["sigma"])
    np.testing.assert_array_equal(a["embeddings"], b["embeddings"])
