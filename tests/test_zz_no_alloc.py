"""scxpca_project_backward_noalloc: the allocation-free projection backward
(dW = (X~ - 1 mu^T)^T dZ with every scratch buffer in the library's fixed
per-device array) against scxpca_project_backward on the same inputs, through
the C ABI. Same kernels, same reduction order: bit-identical outputs. Two
calls in a row share the scratch array, the second with a smaller input, and
the first call on a matrix without a gene order stages one through the array.
"""
from __future__ import annotations

import sys

import numpy as np
import pytest
import scipy.sparse as sp

from scx_amd import synth

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _device_matrix(gpu, m, **kw):
    import torch

    dev = torch.device("cuda:0")
    t = (torch.from_numpy(m.indptr.astype(np.int64)).to(dev),
         torch.from_numpy(m.indices.astype(np.int32)).to(dev),
         torch.from_numpy(m.data.astype(np.float32)).to(dev))
    torch.cuda.synchronize()
    if not kw:
        kw = dict(target_sum=1e4, log1p=True)
    return gpu.Matrix(*t, m.shape, device=0, **kw), t


def _both(mat, n, nv, k, with_mean, seed):
    """(noalloc dW, reference dW) for one random dZ (and mean)."""
    import torch

    dev = torch.device("cuda:0")
    g = torch.Generator(device="cpu").manual_seed(seed)
    dz = torch.randn(n, k, generator=g).to(dev)
    mean = (torch.rand(nv, generator=g) * 0.5).to(dev) if with_mean else None
    dw_na = torch.full((nv, k), float("nan"), device=dev)
    dw_ref = torch.full((nv, k), float("nan"), device=dev)
    torch.cuda.synchronize()
    mat.project_backward_noalloc(dz, k, mean, dw_na)
    mat.project_backward(dz, k, mean, dw_ref)
    torch.cuda.synchronize()
    return dw_na.cpu().numpy(), dw_ref.cpu().numpy()


def test_noalloc_backward_matches_backward_two_calls(gpu):
    big = synth.generate(synth.CFG1, n_obs=4096, workers=1).to_scipy().tocsr()
    small = big[:700, :1500].tocsr()
    small.sort_indices()
    mb, keep_b = _device_matrix(gpu, big)
    ms, keep_s = _device_matrix(gpu, small)
    try:
        # first call on a fresh matrix: no gene order yet (staged in the array)
        say("case: noalloc 4096x2000 k=64 centred (first call, order staged)")
        na, ref = _both(mb, 4096, 2000, 64, True, 1)
        assert np.all(np.isfinite(ref))
        np.testing.assert_array_equal(na, ref)
        say("case: noalloc 700x1500 k=20 uncentred (smaller, packed dZ / dW)")
        na, ref = _both(ms, 700, 1500, 20, False, 2)
        np.testing.assert_array_equal(na, ref)
        say("case: noalloc 4096x2000 k=100 centred (KP=128, order from the matrix)")
        na, ref = _both(mb, 4096, 2000, 100, True, 3)
        np.testing.assert_array_equal(na, ref)
        say("case: noalloc 700x1500 k=64 centred (smaller again)")
        na, ref = _both(ms, 700, 1500, 64, True, 4)
        np.testing.assert_array_equal(na, ref)
    finally:
        mb.close()
        ms.close()


def test_noalloc_backward_pflog_matrix(gpu):
    """A PFlog matrix adds 1 (b^T dZ) through the same scratch array."""
    x = synth.generate(synth.CFG1, n_obs=1000, workers=1).to_scipy().tocsr()
    m, keep = _device_matrix(gpu, x, pflog={"alpha": 0.3, "delta_form": 0})
    try:
        say("case: noalloc PFlog 1000x2000 k=64")
        na, ref = _both(m, 1000, 2000, 64, True, 5)
        np.testing.assert_array_equal(na, ref)
    finally:
        m.close()


def test_noalloc_backward_rejects_oversize(gpu):
    """A matrix whose partials exceed the fixed 64 MiB array is rejected with
    the size in the message (40,000 genes: 8 tile ranges x 40,000 x 64 f32)."""
    import torch

    rng = np.random.default_rng(0)
    x = sp.random(300, 40000, density=0.01, random_state=rng, format="csr", dtype=np.float32)
    x.data[:] = rng.integers(1, 5, size=x.nnz)
    x.sort_indices()
    m, keep = _device_matrix(gpu, x)
    try:
        dev = torch.device("cuda:0")
        dz = torch.randn(300, 64).to(dev)
        dw = torch.empty(40000, 64, device=dev)
        say("case: noalloc oversize 300x40000 k=64")
        with pytest.raises(ValueError, match="bytes of scratch, above the entry point's fixed 67108864 bytes"):
            m.project_backward_noalloc(dz, 64, None, dw)
        m.project_backward(dz, 64, None, dw)  # the allocating entry point still serves it
        torch.cuda.synchronize()
    finally:
        m.close()
