"""Benchmark: one randomized sparse PCA on resident HBM data per step.

Workload (default): 1,000,000 cells x 2,000 genes CSR (~8% density, NB counts,
seed 0), normalize_total(1e4) -> log1p -> implicit centering, k=50, p=10,
3 power iterations, FP32 form, CholeskyQR2. A step is one full rPCA on the
device-resident CSR: row factors, chunk offsets, column statistics, Ω, the
2+2q SpMM passes, the CholeskyQR steps and the Rayleigh-Ritz SVD, with the
embeddings and components written to HBM.

Prints one JSON line. `value` is wall seconds per rPCA (lower is better);
`roofline` reports the dominant SpMM kernel's algorithmic HBM bytes per launch
divided by its average launch time (HIP events on the library stream);
`roofline.traffic` is the measured HBM bytes per launch of that kernel, read
from profiles/spmm_traffic.json (rocprofv3 FETCH_SIZE + WRITE_SIZE, scaled by
the same counters on a plain copy of known size), when that file names the
kernel. `cpu_baseline` times the C/OpenMP restatement of the reference CPU path
(the oracle) on the host cores, on the full matrix by default (--cpu-rows N
takes the first N rows and scales linearly).
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "rPCA wall-s + SpMM HBM GB/s (% peak), 1M cells×2k genes k=50, ≤1e-4 err vs FP64"
HBM_PEAK_GBS = 8000.0


def spmm_bytes(nnz, n_obs, n_vars, k, bf16=False):
    """Algorithmic HBM bytes of one SpMM pass: nonzeros (u32 index + f32 value,
    or bf16 value), indptr, the f64 row factor, the n_obs x k panel written
    (forward) or read (transpose), and the n_vars x k panel."""
    val = 2 if bf16 else 4
    row_factor = 0 if bf16 else 8 * n_obs
    return nnz * (4 + val) + (n_obs + 1) * 8 + row_factor + n_obs * k * 4 + n_vars * k * 4


PROJ_WORKLOAD = "projection_4096x2000_b250"


def run_projection(args):
    """Differentiation form: Z = (X~ - 1 mu^T) W and dW = (X~ - 1 mu^T)^T dZ
    over 4,096-cell x 2,000-gene minibatches of the 10k x 2k generator (seed =
    batch id), W 2,000 x 64 ~ N(0, 1/2,000), dZ ~ N(0, 1). All batches are
    resident in HBM before timing; per batch the timed work is preparing the
    matrix (row factors, X~, chunk offsets, tile-CSC), the forward and the
    backward through torch.autograd."""
    import concurrent.futures as cf

    import numpy as np
    import torch

    from scx_amd import _lib, synth
    from scx_amd.projection import project

    n_b = args.warmup + args.steps
    t0 = time.time()
    with cf.ProcessPoolExecutor(max_workers=min(16, os.cpu_count() or 1)) as ex:
        batches = list(ex.map(_gen_batch, range(n_b)))
    print(f"# generated {n_b} batches in {time.time() - t0:.1f}s", file=sys.stderr, flush=True)
    dev = torch.device("cuda:0")
    res = [tuple(torch.from_numpy(a).to(dev) for a in b) for b in batches]
    nv = 2000
    k = 64
    g = torch.Generator(device="cpu").manual_seed(0)
    w = (torch.randn(nv, k, generator=g) / np.sqrt(nv)).to(dev).requires_grad_(True)
    dz = torch.randn(4096, k, generator=g).to(dev)
    m0 = _lib.Matrix(*res[0], (4096, nv), target_sum=1e4, log1p=True, device=0)
    mean = torch.from_numpy((m0.stats()[0] / 4096).astype(np.float32)).to(dev)
    m0.close()
    torch.cuda.synchronize()
    t_prep = t_fwd = t_bwd = 0.0
    for i in range(n_b):
        timed = i >= args.warmup
        t0 = time.perf_counter()
        m = _lib.Matrix(*res[i], (4096, nv), target_sum=1e4, log1p=True, device=0)
        t1 = time.perf_counter()
        e0, e1, e2 = (torch.cuda.Event(enable_timing=True) for _ in range(3))
        e0.record()
        z = project(m, w, mean)
        e1.record()
        z.backward(dz)
        e2.record()
        torch.cuda.synchronize()
        w.grad = None
        m.close()
        if timed:
            t_prep += t1 - t0
            t_fwd += e0.elapsed_time(e1) * 1e-3
            t_bwd += e1.elapsed_time(e2) * 1e-3
    steps = args.steps
    out = {
        "metric": "projection fwd+bwd wall-s per 4,096-cell x 2,000-gene minibatch (autograd, k=64)",
        "value": round((t_prep + t_fwd + t_bwd) / steps, 6),
        "unit": "s",
        "n_gpus": 1, "steps": steps, "warmup": args.warmup,
        "ms_per_step": round((t_prep + t_fwd + t_bwd) / steps * 1e3, 4),
        "higher_is_better": False, "scaling": "weak", "vs_baseline": None, "dtype": "f32", "data": "synthetic",
        "config": {"workload": f"{PROJ_WORKLOAD}: {steps} timed minibatches of 4,096 x 2,000 (10k x 2k "
                               "generator, seed = batch id), W 2,000 x 64, centered",
                   "parallelism": "single GPU"},
        "ms_prepare": round(t_prep / steps * 1e3, 4),
        "ms_forward": round(t_fwd / steps * 1e3, 4),
        "ms_backward": round(t_bwd / steps * 1e3, 4),
        "backward_over_forward": round(t_bwd / max(t_fwd, 1e-12), 3),
    }
    print(json.dumps(out), flush=True)


PFLOG_METRIC = ("PFlog rPCA wall-s (alpha estimate + delta/baseline + baseline-aware randomized PCA), "
                "1M cells×2k genes k=50, ≤1e-4 err vs FP64")
STREAM_METRIC = ("streamed rPCA wall-s (CSR in host memory, moved over PCIe in shards on every pass), "
                 "1M cells×2k genes k=50, ≤1e-4 err vs FP64")
COV_METRIC = "covariance-route PCA wall-s (exact, method='auto' at <= 5,000 genes), 1M cells×2k genes k=50"
MFMA_F32_PEAK_TFLOPS = 157.3
BF16_MFMA_PEAK_TFLOPS = 2500.0  # MI355X dense bf16 MFMA


def run_covariance(args, x, w, m, dev):
    """The covariance route on the resident matrix: per step refresh + exact
    PCA (Gram, f64 EVD, embeddings). Roofline on the Gram kernel against FP32
    MFMA: dense-equivalent flops n * nv * (nv + 1) (the symmetric half of
    X~^T X~) over its mean launch time (HIP events, timed region). CPU
    baseline: the oracle's C/OpenMP covariance route on two row samples,
    extrapolated linearly in the row count."""
    import torch

    n, nv = x.shape
    emb = torch.empty((n, w.n_comps), dtype=torch.float32, device=dev)
    comps = torch.empty((w.n_comps, nv), dtype=torch.float32, device=dev)

    def step():
        m.refresh()
        return m.pca_covariance(n_comps=w.n_comps, out={"embeddings": emb, "components": comps})

    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()
    m.profile(enable=True, stages=["gram"])
    t0 = time.perf_counter()
    res = None
    for _ in range(args.steps):
        res = step()
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    prof = m.profile(enable=False)
    ms_per_step = (t1 - t0) * 1e3 / args.steps
    m.profile(enable=True)
    for _ in range(args.profile_steps):
        step()
    torch.cuda.synchronize()
    stage_prof = m.profile(enable=False)
    gram_ms = prof["gram"][0] / max(prof["gram"][1], 1)
    flops = float(n) * nv * (nv + 1)
    achieved = flops / (gram_ms * 1e-3) / 1e12
    cpu = None
    if args.cpu_rows != 0:
        import oracle

        cores = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else os.cpu_count()
        cores = min(cores, 16)
        r1 = min(50_000, n)
        r2 = min(100_000, n)
        ts = []
        for rows in (r1, r2):
            xs = x.head(rows)
            tc = time.perf_counter()
            oracle.covariance_pca(oracle.CsrView(xs.indptr, xs.indices, xs.data, xs.shape), n_comps=w.n_comps,
                                  target_sum=w.target_sum, log1p=True, nthreads=cores)
            ts.append(time.perf_counter() - tc)
        slope = (ts[1] - ts[0]) / max(r2 - r1, 1)
        est = ts[1] + slope * (n - r2)
        cpu = {"value": round(est, 3), "unit": "s", "cores": cores, "kind": "port",
               "sample": f"first {r1:,} and {r2:,} of {n:,} rows ({ts[0]:.2f}s, {ts[1]:.2f}s), extrapolated "
                         f"linearly in the row count; the C/OpenMP restatement of accumulate_covariance_into "
                         f"(f64) with LAPACK syevd and the oracle's SpMM, {cores} threads"}
    out = {
        "metric": COV_METRIC,
        "value": round(ms_per_step / 1e3, 6),
        "unit": "s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 4),
        "higher_is_better": False,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32 Gram (f64 across tiles) + f64 EVD",
        "data": "synthetic",
        "config": {
            "workload": f"{w.name}: {n:,} x {nv:,} CSR, nnz={x.nnz:,}, NB counts (seed {w.seed}), "
                        f"normalize_total(1e4)+log1p+centering, k={w.n_comps}, covariance route",
            "n_obs": n, "n_vars": nv, "nnz": x.nnz, "k": w.n_comps, "parallelism": "single GPU",
        },
        "roofline": {
            "bound": "mfma",
            "kernel": "gram stage (k_cov_gram + k_cov_finalize)",
            "achieved": round(achieved, 2),
            "peak": MFMA_F32_PEAK_TFLOPS,
            "unit": "TFLOP/s",
            "frac": round(achieved / MFMA_F32_PEAK_TFLOPS, 4),
            "traffic": None,
            "avg_launch_ms": round(gram_ms, 4),
            "flops_per_launch": flops,
        },
        "stage_ms_per_step": {s: round(v[0] / max(args.profile_steps, 1), 4) for s, v in stage_prof.items() if v[1]},
        "cpu_baseline": cpu,
        "seconds_device_last": round(res["seconds_device"], 6),
    }
    print(json.dumps(out), flush=True)


DECODE_METRIC = "SCX shard decode wall-s, host shard bytes -> device CSR (i64/i32/f32), Scx1 shards of 16,384 rows"


def run_decode(args, x, w):
    """--method decode: the workload's CSR written as Scx1 shards (16,384 rows,
    framed at 256 rows: the reference writer's defaults) by the library's
    writer, then decoded onto the GPU. A step is one decode of every shard,
    from host bytes to the device CSR (host walks + decompression, the upload
    and the kernels); stats split it."""
    import numpy as np
    import torch

    from scx_amd import shards as host

    vals = x.data.astype(np.uint16)
    if not np.array_equal(vals.astype(np.float32), x.data):
        raise SystemExit("decode bench: the synthetic counts do not fit u16")
    t0 = time.time()
    shards = host.encode_csr_shards(x.indptr, x.indices, vals, x.shape[1], rows_per_shard=16384, codec="scx1",
                                    row_group_rows=256)
    enc_bytes = sum(len(s) for s in shards)
    print(f"# encoded {len(shards)} shards, {enc_bytes / 2**20:.1f} MiB in {time.time() - t0:.1f}s", file=sys.stderr,
          flush=True)
    csr = None
    for _ in range(args.warmup):
        csr = host.decode_csr_shards_to_device(shards)
    torch.cuda.synchronize()
    stats = []
    t0 = time.perf_counter()
    for _ in range(args.steps):
        csr = host.decode_csr_shards_to_device(shards)
        stats.append(csr.stats)
    torch.cuda.synchronize()
    ms = (time.perf_counter() - t0) * 1e3 / args.steps
    ok = (np.array_equal(csr.indptr.cpu().numpy(), x.indptr) and np.array_equal(csr.indices.cpu().numpy(),
                                                                                x.indices.view(np.int32))
          and np.array_equal(csr.data.cpu().numpy(), x.data))
    # the path's consumer: the headline rPCA on the decoded device CSR, and the
    # whole chain from shard bytes (decode + resident matrix + rPCA)
    from scx_amd import _lib

    kw = dict(n_comps=w.n_comps, n_oversamples=w.n_oversamples, n_power_iterations=w.n_power_iterations, seed=0,
              qr_method=_lib.QR_CHOLESKY)
    chain = []
    for it in range(3):
        torch.cuda.synchronize()
        tc = time.perf_counter()
        c2 = host.decode_csr_shards_to_device(shards)
        mm = _lib.Matrix(c2.indptr, c2.indices, c2.data, c2.shape, target_sum=w.target_sum, log1p=True)
        mm.pca(**kw)
        torch.cuda.synchronize()
        chain.append(time.perf_counter() - tc)
        mm.close()
        del c2
    med = lambda k: float(np.median([s[k] for s in stats]))  # noqa: E731
    dev_s = med("device_seconds")
    n, nnz = x.shape[0], x.nnz
    # algorithmic device bytes: the staged payload read once, the tables, the
    # decoded indices and values written once
    payload = stats[-1]["host_uploaded_bytes"] - (n + 1) * 8
    moved = payload + nnz * 8
    out_bytes = (n + 1) * 8 + nnz * 8
    cpu = None
    if args.cpu_rows != 0:
        from concurrent.futures import ThreadPoolExecutor

        from oracle import shards as S

        cores = min(len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else os.cpu_count(), 16)

        def cpu_decode():
            with ThreadPoolExecutor(cores) as ex:
                list(ex.map(S.decode_shard, shards))

        med_c, runs = _median_runs(cpu_decode, max(1, min(args.cpu_runs, 3)))
        cpu = {"value": round(med_c, 4), "unit": "s", "cores": cores, "kind": "port",
               "sample": f"all {len(shards)} shards, 1 warm-up + median of {len(runs)}; the oracle's bit-serial "
                         f"C restatement of decode_shard_scipy, one shard per thread",
               "runs_s": [round(t, 4) for t in runs], "cpu_model": _cpu_model()}
    out = {
        "metric": DECODE_METRIC,
        "value": round(ms / 1e3, 6),
        "unit": "s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms, 4),
        "higher_is_better": False,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u32",
        "data": "synthetic",
        "config": {"workload": f"{w.name}: {n:,} x {x.shape[1]:,} CSR, nnz={nnz:,} as {len(shards)} Scx1 shards "
                               f"(16,384 rows, framed at 256 rows, u16 indices and values), "
                               f"{enc_bytes / 2**20:.1f} MiB encoded",
                   "n_obs": n, "n_vars": x.shape[1], "nnz": nnz, "parallelism": "single GPU"},
        "roofline": {"bound": "hbm", "kernel": "k_forbp_rows + k_rice_blocks", "achieved": round(moved / dev_s / 1e9, 1),
                     "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": round(moved / dev_s / 1e9 / HBM_PEAK_GBS, 4),
                     "traffic": None, "avg_launch_ms": round(dev_s * 1e3, 4), "bytes_per_launch": moved},
        "decoded_gbs_end_to_end": round(out_bytes / (ms * 1e-3) / 1e9, 2),
        "stage_ms": {"host": round(med("host_seconds") * 1e3, 3), "upload": round(med("upload_seconds") * 1e3, 3),
                     "device": round(dev_s * 1e3, 3), "total": round(med("total_seconds") * 1e3, 3)},
        "shards_to_pca_ms": round(float(np.median(chain)) * 1e3, 3),
        "shards_to_pca_note": "decode + resident matrix (prepare) + rPCA (k=50, p=10, q=3, FP32, cholesky), "
                              "sigma / embeddings / components copied to host; median of 3",
        "host_uploaded_bytes": stats[-1]["host_uploaded_bytes"],
        "round_trip_exact": bool(ok),
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)


def run_streamed(args, x, w):
    """The streamed operator on the host CSR: every pass of the power loop moves
    the CSR to HBM in shards of --shard-rows rows (pinned staging, a copy stream
    overlapping the next shard's transfer). The value is PCIe-inclusive by
    construction: the matrix never stays in HBM."""
    import numpy as np

    from scx_amd import _lib

    n, nv = x.shape
    qr = _lib.QR_CHOLESKY if args.qr == "cholesky" else _lib.QR_HOUSEHOLDER
    kw = dict(n_comps=w.n_comps, n_oversamples=w.n_oversamples, n_power_iterations=w.n_power_iterations, seed=0,
              qr_method=qr, target_sum=w.target_sum, log1p=True)

    ix, dv = x.indices, x.data
    if args.pinned_host:
        import torch

        pix = torch.from_numpy(x.indices.view(np.int32)).pin_memory()
        pdv = torch.from_numpy(x.data).pin_memory()
        ix, dv = pix.numpy().view(np.uint32), pdv.numpy()

    def step():
        return _lib.randomized_pca_streamed(x.indptr, ix, dv, x.shape, args.shard_rows, **kw)

    for _ in range(args.warmup):
        step()
    t0 = time.perf_counter()
    res = None
    for _ in range(args.steps):
        res = step()
    t1 = time.perf_counter()
    s_per = (t1 - t0) / args.steps
    # passes over the CSR: statistics, 1 + q forward, q + 1 transposed
    passes = 1 + 2 * (w.n_power_iterations + 1)
    csr_bytes = 8 * (n + 1) + 8 * x.nnz
    moved = passes * csr_bytes
    shards = (n + args.shard_rows - 1) // args.shard_rows
    cpu = None
    if args.cpu_rows != 0:
        import oracle

        rows = n if args.cpu_rows < 0 else min(args.cpu_rows, n)
        xs = x.head(rows)
        cores = min(len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else os.cpu_count(), 16)
        tc = time.perf_counter()
        oracle.randomized_pca(oracle.CsrView(xs.indptr, xs.indices, xs.data, xs.shape), n_comps=w.n_comps,
                              n_oversamples=w.n_oversamples, n_power_iterations=w.n_power_iterations, seed=0,
                              target_sum=w.target_sum, log1p=True, nthreads=cores)
        tcpu = time.perf_counter() - tc
        cpu = {"value": round(tcpu * n / rows, 3), "unit": "s", "cores": cores, "kind": "port",
               "sample": (f"the full {n:,}-row matrix" if rows == n else
                          f"first {rows:,} of {n:,} rows, {tcpu:.2f}s, scaled linearly")
                         + f"; one run of the C/OpenMP restatement of randomized_pca_inmemory (f64), {cores} threads"}
    out = {
        "metric": STREAM_METRIC,
        "value": round(s_per, 6),
        "unit": "s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(s_per * 1e3, 3),
        "higher_is_better": False,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {
            "workload": f"{w.name}: {n:,} x {nv:,} CSR in host memory, nnz={x.nnz:,}, NB counts (seed {w.seed}), "
                        f"normalize_total(1e4)+log1p+centering, k={w.n_comps}, p={w.n_oversamples}, "
                        f"q={w.n_power_iterations}, fp32, qr={args.qr}, {shards} shards of {args.shard_rows:,} rows, "
                        + ("page-locked" if args.pinned_host else "pageable") + " host arrays",
            "n_obs": n, "n_vars": nv, "nnz": x.nnz, "shard_rows": args.shard_rows, "parallelism": "single GPU",
        },
        "host_link": {
            "bytes_per_call": moved, "passes": passes, "achieved_GBps": round(moved / s_per / 1e9, 2),
            "note": "CSR bytes moved host -> HBM per call (indptr + indices + values per pass) / wall time",
        },
        "sigma_top3": [round(float(v), 6) for v in res["sigma"][:3]],
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)


def _cpu_model() -> str:
    """The host CPU model (lscpu's 'Model name', read from /proc/cpuinfo)."""
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def _median_runs(fn, runs):
    fn()  # warm-up
    ts = []
    for _ in range(max(runs, 1)):
        tc = time.perf_counter()
        fn()
        ts.append(time.perf_counter() - tc)
    ts.sort()
    return ts[len(ts) // 2], ts


def cpu_baseline_rpca(args, x, w, pflog=False):
    """The oracle (the C/OpenMP restatement of the reference CPU path, f64)
    on row samples of the same matrix: all host cores (at most 16) and one
    thread, one warm-up and the median of --cpu-runs runs each, scaled
    linearly in the row count to the full matrix."""
    import oracle

    n = x.shape[0]
    cores = min(len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else os.cpu_count(), 16)

    def run(rows, threads):
        xs = x.head(rows)
        xv = oracle.CsrView(xs.indptr, xs.indices, xs.data, xs.shape)
        if pflog:
            return lambda: oracle.pflog_pca(xv, n_comps=w.n_comps, n_oversamples=w.n_oversamples,
                                            n_power_iterations=w.n_power_iterations, seed=0, alpha=None, form=1,
                                            nthreads=threads)
        return lambda: oracle.randomized_pca(xv, n_comps=w.n_comps, n_oversamples=w.n_oversamples,
                                             n_power_iterations=w.n_power_iterations, seed=0,
                                             target_sum=w.target_sum, log1p=True, nthreads=threads)

    # the whole matrix up to 250,000 rows, else a quarter of it (at most 250,000 rows)
    rows = (n if n <= 250_000 else min(250_000, n // 4)) if args.cpu_rows < 0 else min(args.cpu_rows, n)
    med, ts = _median_runs(run(rows, cores), args.cpu_runs)
    what = ("alpha estimate, delta, baseline and pflog_pca" if pflog else "randomized_pca_inmemory")
    out = {"value": round(med * n / rows, 3), "unit": "s", "cores": cores, "kind": "port",
           "sample": f"first {rows:,} of {n:,} rows; 1 warm-up + median of {len(ts)} runs ({med:.3f} s), scaled "
                     f"linearly to {n:,} rows; the C/OpenMP restatement of {what} (f64), {cores} threads",
           "runs_s": [round(t, 3) for t in ts], "cpu_model": _cpu_model()}
    if args.cpu_rows_1t != 0:
        rows1 = (n if n <= 25_000 else min(25_000, n // 40)) if args.cpu_rows_1t < 0 else min(args.cpu_rows_1t, n)
        med1, ts1 = _median_runs(run(rows1, 1), args.cpu_runs)
        out["one_thread"] = {"value": round(med1 * n / rows1, 3), "unit": "s", "cores": 1,
                             "sample": f"first {rows1:,} rows; 1 warm-up + median of {len(ts1)} runs ({med1:.3f} s), "
                                       f"scaled linearly to {n:,} rows",
                             "runs_s": [round(t, 3) for t in ts1]}
    return out


def accuracy_block(x, w, qr, rows=50_000):
    """Accuracy of the GPU result on the first `rows` cells (the bench's
    randomized route, FP32 and BF16 sketch): singular values against the
    oracle's f64 randomized PCA on the same Omega (the path's <= 1e-4
    contract) and against the exact PCA of the explicitly centred X~ (dense
    FP64, LAPACK eigendecomposition of the Gram matrix): the algorithm's own
    approximation error, which the f64 randomized PCA shares."""
    import numpy as np

    import oracle
    from scx_amd import _lib

    n, nv = x.shape
    rows = min(rows, n)
    if nv > 5000:
        return None
    xs = x.head(rows)
    view = oracle.CsrView(xs.indptr, xs.indices, xs.data, xs.shape)
    ref = oracle.randomized_pca(view, n_comps=w.n_comps, n_oversamples=w.n_oversamples,
                                n_power_iterations=w.n_power_iterations, seed=0, target_sum=w.target_sum,
                                log1p=True, nthreads=min(16, os.cpu_count() or 1))
    xt = oracle.transform(view, w.target_sum, True)
    d = np.zeros((rows, nv))
    d[np.repeat(np.arange(rows), np.diff(xs.indptr)), xs.indices.astype(np.int64)] = xt
    d -= d.mean(axis=0)
    ev = np.linalg.eigvalsh(d.T @ d)[::-1]
    s_exact = np.sqrt(np.maximum(ev, 0.0))
    del d
    m = _lib.Matrix(xs.indptr, xs.indices, xs.data, xs.shape, target_sum=w.target_sum, log1p=True)
    out = {"cells": rows, "reference": "oracle f64 rPCA (same Omega) and dense FP64 exact PCA of the subsample"}
    try:
        for name, prec in (("fp32", _lib.PRECISION_FP32), ("bf16", _lib.PRECISION_BF16_SKETCH)):
            got = m.pca(n_comps=w.n_comps, n_oversamples=w.n_oversamples, n_power_iterations=w.n_power_iterations,
                        seed=0, qr_method=qr, precision=prec)
            nc = w.n_comps
            e64 = np.abs(got["sigma"][:nc] - ref["sigma_all"][:nc]) / ref["sigma_all"][:nc]
            eex = np.abs(got["sigma"][:nc] - s_exact[:nc]) / s_exact[:nc]
            out[name] = {"sigma_rel_vs_f64_rpca_max": float(f"{e64.max():.3g}"),
                         "sigma_rel_vs_dense_fp64_top1": float(f"{eex[0]:.3g}"),
                         "sigma_rel_vs_dense_fp64_top10_max": float(f"{eex[:10].max():.3g}"),
                         "sigma_rel_vs_dense_fp64_max": float(f"{eex.max():.3g}")}
        e_alg = np.abs(ref["sigma_all"][:w.n_comps] - s_exact[:w.n_comps]) / s_exact[:w.n_comps]
        out["f64_rpca_sigma_rel_vs_dense_fp64_max"] = float(f"{e_alg.max():.3g}")
    finally:
        m.close()
    return out


def _gen_batch(b):
    from scx_amd import synth

    x = synth.generate(synth.CFG1, n_obs=4096, seed=b, workers=1)
    return x.indptr, x.indices.view("int32"), x.data


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--workload", default="pca_1m_x_2k_k50")
    ap.add_argument("--precision", choices=["fp32", "bf16"], default="fp32")
    ap.add_argument("--qr", choices=["cholesky", "householder"], default="cholesky")
    ap.add_argument("--cpu-rows", type=int, default=-1,
                    help="rows of the all-cores CPU baseline sample: -1 the whole matrix up to 250,000 rows, "
                         "else a quarter of it (at most 250,000 rows); 0 skip the CPU baseline; N the first N rows")
    ap.add_argument("--cpu-rows-1t", type=int, default=-1,
                    help="rows of the one-thread CPU baseline sample: -1 the whole matrix up to 25,000 rows, "
                         "else a fortieth of it (at most 25,000 rows); 0 skip")
    ap.add_argument("--cpu-runs", type=int, default=5,
                    help="timed CPU baseline runs after one warm-up (the median is reported)")
    ap.add_argument("--legs", type=int, default=5, metavar="STEPS",
                    help="steps of the extra timed legs (qr=householder, the BF16 sketch form); 0 skips them")
    ap.add_argument("--host-input", type=int, default=3, metavar="RUNS",
                    help="also time RUNS calls from host CSR buffers (H2D copy, validation, prepare, rPCA, "
                         "results back to host); reported beside the device-resident value, never as it")
    ap.add_argument("--shard-rows", type=int, default=250_000,
                    help="rows per shard of --method streamed")
    ap.add_argument("--pinned-host", action="store_true",
                    help="--method streamed: hold the host CSR's indices and values in page-locked memory "
                         "(sent to the device directly instead of through the pinned staging copy)")
    ap.add_argument("--method", choices=["randomized", "covariance", "pflog", "streamed", "decode"],
                    default="randomized",
                    help="randomized (the headline metric), the exact covariance route, PFlog (v4) "
                         "normalization with its baseline-aware randomized PCA on the raw counts, or the "
                         "streamed operator for a CSR larger than HBM (host CSR, shards over PCIe each pass)")
    ap.add_argument("--profile-steps", type=int, default=2,
                    help="untimed calls after the timed ones, every stage timed, for stage_ms_per_step")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="before the timed region, copy MIB MiB device-to-device once (a plain copy "
                         "of known size whose FETCH_SIZE/WRITE_SIZE calibrate the traffic counters)")
    args = ap.parse_args()

    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print(json.dumps({"error": "this build runs on one GPU only; --gpus must be 1 "
                                   "(the multi-GPU row sharding is a design note, not implemented)"}))
        sys.exit(2)
    if args.workload == PROJ_WORKLOAD:
        if "--steps" not in sys.argv:
            args.steps = 250
        if "--warmup" not in sys.argv:
            args.warmup = 5
        run_projection(args)
        return

    import numpy as np
    import torch

    from scx_amd import _lib, synth

    w = synth.WORKLOADS[args.workload]
    t0 = time.time()
    x = synth.generate(w)
    gen_s = time.time() - t0
    n, nv = x.shape
    k = min(w.n_comps + w.n_oversamples, n, nv)
    print(f"# generated {n}x{nv} nnz={x.nnz} in {gen_s:.1f}s", file=sys.stderr, flush=True)

    if args.method == "streamed":
        run_streamed(args, x, w)
        return
    if args.method == "decode":
        run_decode(args, x, w)
        return

    dev = torch.device("cuda:0")
    d_indptr = torch.from_numpy(x.indptr).to(dev)
    d_idx = torch.from_numpy(x.indices.view(np.int32)).to(dev)
    d_val = torch.from_numpy(x.data).to(dev)
    emb = torch.empty((n, w.n_comps), dtype=torch.float32, device=dev)
    comps = torch.empty((w.n_comps, nv), dtype=torch.float32, device=dev)
    torch.cuda.synchronize()

    pflog = args.method == "pflog"
    pf = {"alpha": None, "delta_form": _lib.PFLOG_DELTA_F64} if pflog else None
    m = _lib.Matrix(d_indptr, d_idx, d_val, x.shape, target_sum=w.target_sum, log1p=True, device=0, pflog=pf)
    if args.method == "covariance":
        run_covariance(args, x, w, m, dev)
        m.close()
        return
    qr = _lib.QR_CHOLESKY if args.qr == "cholesky" else _lib.QR_HOUSEHOLDER
    prec = _lib.PRECISION_BF16_SKETCH if args.precision == "bf16" else _lib.PRECISION_FP32
    kw = dict(n_comps=w.n_comps, n_oversamples=w.n_oversamples, n_power_iterations=w.n_power_iterations,
              seed=0, qr_method=qr, precision=prec, out={"embeddings": emb, "components": comps})

    if args.copy_calibration > 0:
        nb = args.copy_calibration << 20
        src = torch.ones(nb // 4, dtype=torch.float32, device=dev)
        dst = torch.empty_like(src)
        dst.copy_(src)
        torch.cuda.synchronize()
        print(f"# copy calibration: {nb} bytes read, {nb} bytes written", file=sys.stderr, flush=True)
        del src, dst

    def step():
        m.refresh()
        return m.pca(**kw)

    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()
    # timed region: HIP events only around the SpMM launches (each timed stage
    # costs two event records on the stream)
    bf16 = args.precision == "bf16"
    cands = ["spmm_fwd", "spmm_t"] + (["spmm_fwd_bf16", "spmm_t_bf16"] if bf16 else [])
    m.profile(enable=True, stages=cands)
    t0 = time.perf_counter()
    res = None
    for _ in range(args.steps):
        res = step()
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    prof = m.profile(enable=False)
    ms_per_step = (t1 - t0) * 1e3 / args.steps
    # stage split: further calls with every stage timed, outside the timed region
    m.profile(enable=True)
    for _ in range(args.profile_steps):
        step()
    torch.cuda.synchronize()
    stage_prof = m.profile(enable=False)

    # dominant SpMM kernel
    dom = max(cands, key=lambda s: prof[s][0])
    dom_ms = prof[dom][0] / max(prof[dom][1], 1)
    bytes_per = spmm_bytes(x.nnz, n, nv, k, bf16=dom.endswith("bf16"))
    achieved = bytes_per / (dom_ms * 1e-3) / 1e9
    traffic = None
    tf = os.path.join(ROOT, "profiles", "spmm_traffic.json")
    if os.path.exists(tf) and not pflog:
        with open(tf) as f:
            rec = json.load(f).get(f"{w.name}:{args.precision}:{dom}")
        if rec:
            traffic = rec["traffic_bytes"]

    # the Gram of the tall CholeskyQR (the one dense contraction) on FP32 MFMA:
    # algorithmic flops 2 n k'^2 per launch over the stage's mean launch time
    gram = None
    if stage_prof.get("gram", (0.0, 0))[1]:
        g_ms = stage_prof["gram"][0] / stage_prof["gram"][1]
        g_flops = 2.0 * n * k * k
        gram = {"launches_per_step": stage_prof["gram"][1] // max(args.profile_steps, 1),
                "avg_ms": round(g_ms, 4), "flops_per_launch": g_flops,
                "achieved_tflops": round(g_flops / (g_ms * 1e-3) / 1e12, 2), "peak_tflops": MFMA_F32_PEAK_TFLOPS,
                "frac": round(g_flops / (g_ms * 1e-3) / 1e12 / MFMA_F32_PEAK_TFLOPS, 4),
                "note": "stage = k_gram4 (16x16x4 f32 MFMA, f64 partials) + k_sum_parts16"}

    # extra legs on the same resident matrix, after the timed region
    legs = None
    if args.legs > 0 and not pflog:
        legs = {}
        for name, kw_leg in (("qr_householder", dict(kw, qr_method=_lib.QR_HOUSEHOLDER)),
                             ("bf16_sketch", dict(kw, precision=_lib.PRECISION_BF16_SKETCH)),
                             ("fp32", dict(kw, precision=_lib.PRECISION_FP32))):
            if (name == "qr_householder" and qr == _lib.QR_HOUSEHOLDER) or (name == "fp32" and not bf16):
                continue
            m.refresh()
            r0 = m.pca(**kw_leg)
            torch.cuda.synchronize()
            tl = time.perf_counter()
            for _ in range(args.legs):
                m.refresh()
                r1 = m.pca(**kw_leg)
            torch.cuda.synchronize()
            ms_leg = (time.perf_counter() - tl) * 1e3 / args.legs
            se = np.abs(r1["sigma"][:w.n_comps] - res["sigma"][:w.n_comps]) / np.abs(res["sigma"][:w.n_comps])
            legs[name] = {"ms_per_step": round(ms_leg, 4), "steps": args.legs,
                          "sigma_rel_vs_headline_max": float(f"{se.max():.3g}")}
            del r0

    cpu = cpu_baseline_rpca(args, x, w, pflog) if args.cpu_rows != 0 else None

    accuracy = None
    if not pflog and args.legs > 0:
        accuracy = accuracy_block(x, w, qr)

    host_input = None
    if args.host_input > 0:
        kw_h = dict(kw)
        kw_h.pop("out")
        times = []
        for _ in range(args.host_input):
            torch.cuda.synchronize()
            th = time.perf_counter()
            mh = _lib.Matrix(x.indptr, x.indices, x.data, x.shape, target_sum=w.target_sum, log1p=True, pflog=pf)
            mh.pca(**kw_h)
            mh.close()
            times.append(time.perf_counter() - th)
        host_input = {"value": round(float(np.median(times)), 4), "unit": "s", "runs": args.host_input,
                      "includes": "host CSR -> HBM copy, row checks, prepare, rPCA, embeddings/components to host"}

    stage_ms = {s: round(v[0] / max(args.profile_steps, 1), 4) for s, v in stage_prof.items() if v[1]}
    out = {
        "metric": PFLOG_METRIC if pflog else METRIC,
        "value": round(ms_per_step / 1e3, 6),
        "unit": "s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 4),
        "higher_is_better": False,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32" if not bf16 else "bf16+f32",
        "data": "synthetic",
        "config": {
            "workload": f"{w.name}: {n:,} x {nv:,} CSR, nnz={x.nnz:,}, NB counts (seed {w.seed}), "
                        + ("PFlog v4 (alpha estimated on the GPU, in-memory delta form) + per-cell baseline "
                           "+ centering" if pflog else "normalize_total(1e4)+log1p+centering")
                        + f", k={w.n_comps}, p={w.n_oversamples}, q={w.n_power_iterations}, {args.precision}, "
                          f"qr={args.qr}",
            "n_obs": n, "n_vars": nv, "nnz": x.nnz, "k": k,
            "parallelism": "single GPU",
        },
        "roofline": {
            "bound": "hbm",
            "kernel": dom,
            "achieved": round(achieved, 1),
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4),
            "traffic": traffic,
            "avg_launch_ms": round(dom_ms, 4),
            "bytes_per_launch": bytes_per,
        },
        "stage_ms_per_step": stage_ms,
        "gram_mfma": gram,
        "legs": legs,
        "accuracy": accuracy if not pflog else None,
        "cpu_baseline": cpu,
        "seconds_device_last": round(res["seconds_device"], 6),
        "host_input": host_input,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
