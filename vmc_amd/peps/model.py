"""PEPS container (mirror of src/vmc/peps/standard/model.py:22-98).

Site tensors are torch tensors on the HIP device with the reference layout
``(d, up, down, left, right)``; the flat parameter vector (ravel_pytree order)
is what the C ABI consumes.
"""
from __future__ import annotations

import numpy as np
import torch

from vmc_amd.peps.strategy import ContractionStrategy, Variational


class PEPS:
    """Open-boundary PEPS on a rectangular lattice."""

    @staticmethod
    def site_dims(row, col, n_rows, n_cols, bond_dim):
        up = 1 if row == 0 else bond_dim
        down = 1 if row == n_rows - 1 else bond_dim
        left = 1 if col == 0 else bond_dim
        right = 1 if col == n_cols - 1 else bond_dim
        return up, down, left, right

    def __init__(self, *, shape, bond_dim, phys_dim=2, contraction_strategy: ContractionStrategy | None = None,
                 dtype=torch.float64, tensors=None, seed=0, device="cuda"):
        self.shape = (int(shape[0]), int(shape[1]))
        self.bond_dim = int(bond_dim)
        self.phys_dim = int(phys_dim)
        self.dtype = dtype
        self.device = torch.device(device)
        # model.py:51-54: the default strategy is Variational(D^2)
        self.strategy = contraction_strategy or Variational(self.bond_dim * self.bond_dim)
        n_rows, n_cols = self.shape
        if tensors is None:
            from vmc_amd.synthetic import random_site_tensors

            tensors = random_site_tensors(self.shape, self.bond_dim, self.phys_dim, seed=seed)
        self.tensors = [
            [torch.as_tensor(np.asarray(tensors[r][c]), dtype=dtype).to(self.device) for c in range(n_cols)]
            for r in range(n_rows)
        ]
        for r in range(n_rows):
            for c in range(n_cols):
                want = (self.phys_dim, *self.site_dims(r, c, n_rows, n_cols, self.bond_dim))
                assert tuple(self.tensors[r][c].shape) == want, (r, c, self.tensors[r][c].shape, want)

    def flat_params(self) -> torch.Tensor:
        return torch.cat([t.reshape(-1) for row in self.tensors for t in row]).contiguous()

    def set_flat_params(self, flat: torch.Tensor) -> None:
        off = 0
        for row in self.tensors:
            for i, t in enumerate(row):
                n = t.numel()
                row[i] = flat[off:off + n].reshape(t.shape).to(t.dtype).clone()
                off += n

    def params_per_site(self):
        n_rows, n_cols = self.shape
        return [int(np.prod(self.site_dims(r, c, n_rows, n_cols, self.bond_dim)))
                for r in range(n_rows) for c in range(n_cols)]

    @staticmethod
    def unflatten_sample(sample, shape):
        return sample.reshape(shape)

    def __call__(self, x):
        """log psi of samples x (n, n_sites) -- PEPS.__call__ (model.py:82-87)."""
        from vmc_amd.peps.compat import log_value

        return log_value(self, x)

    def random_physical_configuration(self, seed: int, n_samples: int = 1):
        rng = np.random.default_rng(seed)
        return torch.as_tensor(
            rng.integers(0, self.phys_dim, size=(n_samples, *self.shape)).astype(np.int32))
