"""Differentiable projection of the centered, normalized matrix:

    Z = (X~ - 1 mu^T) W,         X~ = log1p(normalize_total(X)),

as a ``torch.autograd.Function`` whose backward is hand-written:

    dW  = (X~ - 1 mu^T)^T dZ  = X~^T dZ - mu (1^T dZ)     (the transposed SpMM)
    dmu = -W (1^T dZ)^T                                    (when mu needs a gradient)

Both directions run the PCA's HIP SpMM kernels through the C ABI
(``scxpca_project_forward`` / ``scxpca_project_backward``) on the caller's
current HIP stream; nothing is computed on the CPU. X~ is the prepared
matrix's (computed once when the ``Matrix`` is created or refreshed, from the
raw counts); autograd saves only W, mu and the matrix handle, never a copy of
X~. The reference has no counterpart: this is the differentiation form of the
path (forward Z = X~ W as in ``cpu.rs:1185-1229``, its adjoint as in
``cpu.rs:1238-1267,1282-1423``).
"""
from __future__ import annotations

import torch

from . import _lib


def _check_panel(t: torch.Tensor, rows: int, name: str) -> None:
    if not t.is_cuda or t.dtype != torch.float32 or t.dim() != 2 or t.shape[0] != rows:
        raise ValueError(f"{name} must be a float32 CUDA tensor of shape ({rows}, k)")
    if not 1 <= t.shape[1] <= 128:
        raise ValueError(f"{name} must have 1..128 columns, got {t.shape[1]}")


class _Project(torch.autograd.Function):
    @staticmethod
    def forward(ctx, w: torch.Tensor, matrix: "_lib.Matrix", mean: torch.Tensor | None):
        n, nv = matrix.shape
        w = w.contiguous()
        _check_panel(w, nv, "W")
        if mean is not None:
            mean = mean.contiguous()
            if mean.shape != (nv,) or mean.dtype != torch.float32 or not mean.is_cuda:
                raise ValueError(f"mean must be a float32 CUDA tensor of shape ({nv},)")
        k = w.shape[1]
        z = torch.empty((n, k), dtype=torch.float32, device=w.device)
        stream = torch.cuda.current_stream(w.device).cuda_stream
        matrix.project_forward(w, k, mean, z, stream=stream)
        ctx.matrix = matrix
        ctx.save_for_backward(w, mean if mean is not None else torch.empty(0, device=w.device))
        ctx.has_mean = mean is not None
        return z

    @staticmethod
    def backward(ctx, dz: torch.Tensor):
        w, mean = ctx.saved_tensors
        matrix = ctx.matrix
        n, nv = matrix.shape
        dz = dz.contiguous().to(torch.float32)
        _check_panel(dz, n, "dZ")
        k = dz.shape[1]
        dw = None
        if ctx.needs_input_grad[0]:
            dw = torch.empty((nv, k), dtype=torch.float32, device=dz.device)
            stream = torch.cuda.current_stream(dz.device).cuda_stream
            matrix.project_backward(dz, k, mean if ctx.has_mean else None, dw, stream=stream)
        dmean = None
        if ctx.has_mean and ctx.needs_input_grad[2]:
            dmean = -(w @ dz.sum(dim=0))
        return dw, None, dmean


def project(matrix: "_lib.Matrix", w: torch.Tensor, mean: torch.Tensor | None = None) -> torch.Tensor:
    """Z = (X~ - 1 mean^T) W for a prepared ``Matrix`` (device-resident CSR with
    its normalize_total / log1p settings); differentiable in ``w`` and ``mean``.
    ``w``: float32 CUDA (n_vars, k), k <= 128; ``mean``: float32 CUDA (n_vars,)
    or None (no centering)."""
    return _Project.apply(w, matrix, mean)
