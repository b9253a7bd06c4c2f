"""Copies the reference's golden SCX files used by the shard-decode tests into
tests/golden/scx/, with their expected CSR (the JSON twins the reference's own
tests hold) as compact .npz. Run in a container that has /root/reference:

    python tests/golden/make_scx_fixtures.py

The .scx files are data fixtures (tests/reference_files/golden_*.scx,
v2_multimodal_citeseq.scx); nothing here is code from the reference.
"""
from __future__ import annotations

import json
import os
import shutil

import numpy as np

SRC = "/root/reference/tests/reference_files"
DST = os.path.join(os.path.dirname(os.path.abspath(__file__)), "scx")
GOLDEN = [f"golden_{c}_{v}" for c in ("scx1", "none", "zstd", "lz4shuffle") for v in ("u8", "u16", "u32", "f32")
          if not (c == "scx1" and v == "f32")] + ["golden_pcodec_u16"]


def main():
    os.makedirs(DST, exist_ok=True)
    for name in GOLDEN:
        shutil.copyfile(os.path.join(SRC, name + ".scx"), os.path.join(DST, name + ".scx"))
        js = json.load(open(os.path.join(SRC, name + ".json")))
        np.savez_compressed(os.path.join(DST, name + ".npz"), indptr=np.asarray(js["indptr"], np.int64),
                            indices=np.asarray(js["indices"], np.int32),
                            data=np.asarray(js["data"], np.float64).astype(np.float32),
                            shape=np.asarray([js["n_obs"], js["n_vars"]], np.int64),
                            codec=np.asarray(js["codec"]), value_encoding=np.asarray(js["value_encoding"]))
    name = "v2_multimodal_citeseq"
    shutil.copyfile(os.path.join(SRC, name + ".scx"), os.path.join(DST, name + ".scx"))
    js = json.load(open(os.path.join(SRC, name + ".json")))
    arrays = {}
    for mod in js["expected_csr_per_modality"]:
        i = int(mod["modality_id"])
        arrays[f"m{i}_indptr"] = np.asarray(mod["indptr"], np.int64)
        arrays[f"m{i}_indices"] = np.asarray(mod["indices"], np.int32)
        arrays[f"m{i}_data"] = np.asarray(mod["data"], np.float64).astype(np.float32)
        arrays[f"m{i}_name"] = np.asarray(mod["modality_name"])
    np.savez_compressed(os.path.join(DST, name + ".npz"), **arrays)


if __name__ == "__main__":
    main()
