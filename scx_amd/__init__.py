"""MI355X-native randomized sparse PCA (the scx-accel ``pca`` hot path).

``scx_amd.accel`` mirrors ``pyscx.accel`` (``pca``, ``normalize_total``,
``log1p``); ``scx_amd.projection`` holds the autograd projection; the native
library (HIP kernels + C ABI, ``include/scxpca.h``) is ``scx_amd/csrc``.
"""
from . import accel  # noqa: F401

__version__ = "0.1.0"
