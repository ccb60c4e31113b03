"""Blockade PEPS container (mirror of src/vmc/peps/blockade/model.py:43-264, 1004-1039).

Site tensors (2, nc, mu_u, mu_d, mu_l, mu_r) with nc = 2^(incoming bonds)
and bonds of Dmax = max(D0, D1); samples are int32 occupations (0/1) of the
sites, row-major. The configuration's slice of a site is n * nc + cfg with
cfg = kL * stride + kU for n = 0 and 0 for n = 1 (model.py:157-199); its bond
masks for D0 != D1 (model.py:201-212) are applied by the device library.
"""
from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np
import torch

from vmc_amd.peps.strategy import ContractionStrategy, Variational

__all__ = ["BlockadePEPSConfig", "BlockadePEPS", "random_independent_set"]


@dataclass(frozen=True)
class BlockadePEPSConfig:
    """model.py:43-66."""

    shape: tuple
    D0: int
    D1: int
    phys_dim: int = 2
    dtype: torch.dtype = torch.complex128
    mask_per_charge: np.ndarray | None = field(init=False, repr=False, compare=False)

    def __post_init__(self):
        if self.phys_dim != 2:
            raise ValueError("BlockadePEPS requires phys_dim=2")
        dmax = max(self.D0, self.D1)
        mask = None
        if not (self.D0 == dmax and self.D1 == dmax):
            mask = np.arange(dmax)[None, :] < np.array([self.D0, self.D1])[:, None]
        object.__setattr__(self, "mask_per_charge", mask)

    @property
    def Dmax(self) -> int:
        return max(self.D0, self.D1)


def _can_flip_to_one(config, r, c):
    """model.py:220-239."""
    n_rows, n_cols = config.shape
    return not ((c > 0 and config[r, c - 1] == 1) or (c < n_cols - 1 and config[r, c + 1] == 1)
                or (r > 0 and config[r - 1, c] == 1) or (r < n_rows - 1 and config[r + 1, c] == 1))


def random_independent_set(rng: np.random.Generator, shape) -> np.ndarray:
    """model.py:242-264 with a numpy generator: a greedy row-major pass that
    occupies each site with probability 1/2 where the blockade allows it."""
    config = np.zeros(tuple(shape), dtype=np.int32)
    for r in range(shape[0]):
        for c in range(shape[1]):
            if rng.random() < 0.5 and _can_flip_to_one(config, r, c):
                config[r, c] = 1
    return config


class BlockadePEPS:
    """model.py:69-149."""

    def __init__(self, *, config: BlockadePEPSConfig, contraction_strategy: ContractionStrategy | None = None,
                 tensors=None, seed=0, device="cuda"):
        self.config = config
        self.shape = (int(config.shape[0]), int(config.shape[1]))
        self.D0, self.D1 = int(config.D0), int(config.D1)
        self.Dmax = config.Dmax
        self.bond_dim = self.Dmax
        self.phys_dim = int(config.phys_dim)
        self.dtype = config.dtype
        self.device = torch.device(device)
        self.strategy = contraction_strategy or Variational(self.Dmax * self.Dmax)
        rng = np.random.default_rng(seed)
        self.tensors = []
        for r in range(self.shape[0]):
            row = []
            for c in range(self.shape[1]):
                shp = self.site_shape(r, c)
                if tensors is None:
                    scale = 1.0 / np.sqrt(self.Dmax ** 4 * 2.0)
                    t = rng.standard_normal(shp) * scale
                    if self.dtype.is_complex:
                        t = t + 1j * rng.standard_normal(shp) * scale
                else:
                    t = np.asarray(tensors[r][c])
                    assert t.shape == shp, (r, c, t.shape, shp)
                row.append(torch.as_tensor(t, dtype=self.dtype).to(self.device))
            self.tensors.append(row)

    def nc(self, r, c):
        """2^(incoming bonds): model.py:104-107."""
        return 2 ** (int(r > 0) + int(c > 0))

    def site_shape(self, r, c):
        n_rows, n_cols = self.shape
        dm = self.Dmax
        return (2, self.nc(r, c), dm if r > 0 else 1, dm if r < n_rows - 1 else 1, dm if c > 0 else 1,
                dm if c < n_cols - 1 else 1)

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
        """model.py:1009-1017: parameters of one active (n, cfg) slice."""
        return [int(np.prod(self.site_shape(r, c)[2:])) for r in range(self.shape[0]) for c in range(self.shape[1])]

    def sliced_dims(self):
        """model.py:1031-1039: 2 * nc slices per site."""
        return tuple(2 * self.nc(r, c) for r in range(self.shape[0]) for c in range(self.shape[1]))

    @staticmethod
    def flatten_sample(config) -> torch.Tensor:
        return torch.as_tensor(config).reshape(-1)

    @staticmethod
    def unflatten_sample(sample, shape) -> torch.Tensor:
        return torch.as_tensor(sample).reshape(tuple(shape))

    def random_physical_configuration(self, seed: int, n_samples: int = 1) -> torch.Tensor:
        """model.py:136-147: random independent sets, flattened (n_samples, n_sites)."""
        rng = np.random.default_rng(seed)
        return torch.as_tensor(np.stack([random_independent_set(rng, self.shape).reshape(-1)
                                         for _ in range(n_samples)]))

    def apply(self, sample):
        from vmc_amd.peps.blockade.compat import blockade_apply

        return blockade_apply(self.flat_params(), sample, self.shape, self.config, self.strategy, model=self)
