"""A minimal AnnData stand-in (anndata is not a dependency of this package).

``pca`` and ``pflog`` duck-type their argument: they need ``.X``, ``.var``
(for ``mask_var`` / ``highly_variable``), ``.obs``, ``.obsm``, ``.varm``,
``.layers`` and ``.uns``. A real
``anndata.AnnData`` works unchanged; this class provides the same slots.
"""
from __future__ import annotations

import pandas as pd


class AnnDataLite:
    def __init__(self, X, obs=None, var=None):
        self.X = X
        n_obs, n_vars = X.shape
        self.obs = obs if obs is not None else pd.DataFrame(index=[str(i) for i in range(n_obs)])
        self.var = var if var is not None else pd.DataFrame(index=[f"g{i}" for i in range(n_vars)])
        self.obsm = {}
        self.varm = {}
        self.layers = {}
        self.uns = {}

    @property
    def n_obs(self) -> int:
        return self.X.shape[0]

    @property
    def n_vars(self) -> int:
        return self.X.shape[1]

    @property
    def shape(self):
        return self.X.shape
