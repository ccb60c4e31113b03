"""Gauge-invariant PEPS container (mirror of src/vmc/peps/gi/model.py:40-182, 1046-1080).

Site tensors (phys_dim, nc, mu_u, mu_d, mu_l, mu_r) with nc = N^(links-1)
feasible local link configurations; samples are int32 vectors of sites,
horizontal links and vertical links. A non-uniform degeneracy_per_charge
masks every bond index at or above the degeneracy of its link's charge
(gi/model.py:52-60, 292-305); the device applies the masks per slice.
"""
from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np
import torch

from vmc_amd.peps.strategy import ContractionStrategy, Variational

__all__ = ["GIPEPSConfig", "GIPEPS"]


@dataclass(frozen=True)
class GIPEPSConfig:
    """gi/model.py:40-64."""

    shape: tuple
    N: int
    phys_dim: int
    Qx: int
    degeneracy_per_charge: tuple
    charge_of_site: tuple
    dtype: torch.dtype = torch.complex128
    mask_per_charge: np.ndarray | None = field(init=False, repr=False, compare=False)

    def __post_init__(self):
        deg = tuple(int(d) for d in self.degeneracy_per_charge)
        if len(deg) != int(self.N) or min(deg) < 1:
            raise ValueError("degeneracy_per_charge needs N entries, each at least 1")
        dmax = max(deg)
        mask = None
        if any(d != dmax for d in deg):
            mask = np.arange(dmax)[None, :] < np.asarray(deg)[:, None]
        object.__setattr__(self, "mask_per_charge", mask)

    @property
    def dmax(self) -> int:
        return int(max(self.degeneracy_per_charge))


class GIPEPS:
    """gi/model.py:67-182."""

    def __init__(self, *, config: GIPEPSConfig, contraction_strategy: ContractionStrategy | None = None,
                 tensors=None, seed=0, device="cuda"):
        self.config = config
        self.shape = (int(config.shape[0]), int(config.shape[1]))
        self.N = int(config.N)
        self.phys_dim = int(config.phys_dim)
        self.Qx = int(config.Qx)
        self.charge_of_site = tuple(int(c) % self.N for c in config.charge_of_site)
        self.dmax = config.dmax
        self.bond_dim = self.dmax
        self.dtype = config.dtype
        self.device = torch.device(device)
        self.strategy = contraction_strategy or Variational(self.dmax * self.dmax)
        n_rows, n_cols = self.shape
        rng = np.random.default_rng(seed)
        self.tensors = []
        for r in range(n_rows):
            row = []
            for c in range(n_cols):
                shp = self.site_shape(r, c)
                if tensors is None:
                    scale = 1.0 / np.sqrt(self.dmax ** 4 * self.phys_dim)
                    t = rng.standard_normal(shp) * scale
                    if self.dtype.is_complex:
                        t = t + 1j * rng.standard_normal(shp) * scale
                else:
                    t = np.asarray(tensors[r][c])
                    assert t.shape == shp, (r, c, t.shape, shp)
                row.append(torch.as_tensor(t, dtype=self.dtype).to(self.device))
            self.tensors.append(row)

    def nc(self, r, c):
        n_rows, n_cols = self.shape
        nl = int(r > 0) + int(r < n_rows - 1) + int(c > 0) + int(c < n_cols - 1)
        return int(self.N ** max(nl - 1, 0))

    def site_shape(self, r, c):
        n_rows, n_cols = self.shape
        dm = self.dmax
        return (self.phys_dim, self.nc(r, c), dm if r > 0 else 1, dm if r < n_rows - 1 else 1,
                dm if c > 0 else 1, dm if c < n_cols - 1 else 1)

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
        """gi/model.py:1050-1062: parameters of one active (sigma, cfg) slice."""
        return [int(np.prod(self.site_shape(r, c)[2:])) for r in range(self.shape[0]) for c in range(self.shape[1])]

    def sliced_dims(self):
        """gi/model.py:1065-1080: phys_dim * nc slices per site."""
        return tuple(self.phys_dim * self.nc(r, c) for r in range(self.shape[0]) for c in range(self.shape[1]))

    @staticmethod
    def flatten_sample(sites, h_links, v_links):
        return torch.cat([torch.as_tensor(x).reshape(-1) for x in (sites, h_links, v_links)])

    @staticmethod
    def unflatten_sample(sample, shape):
        n_rows, n_cols = shape
        ns, nh = n_rows * n_cols, n_rows * (n_cols - 1)
        sample = torch.as_tensor(sample)
        return (sample[:ns].reshape(n_rows, n_cols), sample[ns:ns + nh].reshape(n_rows, n_cols - 1),
                sample[ns + nh:].reshape(n_rows - 1, n_cols))

    def random_physical_configuration(self, seed: int, n_samples: int = 1):
        """gi/model.py:153-182 with a numpy generator: links from random plaquette
        deltas on the zero configuration, matter drawn within the Gauss charge."""
        rng = np.random.default_rng(seed)
        n_rows, n_cols = self.shape
        N = self.N
        h = np.zeros((n_samples, n_rows, n_cols - 1), np.int64)
        v = np.zeros((n_samples, n_rows - 1, n_cols), np.int64)
        if n_rows > 1 and n_cols > 1:
            dl = rng.integers(0, N, size=(n_samples, n_rows - 1, n_cols - 1))
            h[:, :n_rows - 1, :] += dl
            h[:, 1:, :] -= dl
            v[:, :, :n_cols - 1] += dl
            v[:, :, 1:] -= dl
            h %= N
            v %= N
        nl = np.pad(h, ((0, 0), (0, 0), (1, 0)))
        nr = np.pad(h, ((0, 0), (0, 0), (0, 1)))
        nu = np.pad(v, ((0, 0), (1, 0), (0, 0)))
        nd = np.pad(v, ((0, 0), (0, 1), (0, 0)))
        charge = (self.Qx - (nl + nd - nu - nr)) % N
        cos = np.asarray(self.charge_of_site)
        sites = np.zeros((n_samples, n_rows, n_cols), np.int64)
        for q in range(N):
            idx = np.nonzero(cos == q)[0]
            mask = charge == q
            if mask.any():
                if len(idx) == 0:
                    raise ValueError("charge_of_site must include all charges 0..N-1")
                sites[mask] = idx[rng.integers(0, len(idx), size=int(mask.sum()))]
        out = np.concatenate([sites.reshape(n_samples, -1), h.reshape(n_samples, -1), v.reshape(n_samples, -1)], 1)
        return torch.as_tensor(out.astype(np.int32))

    def apply(self, sample):
        from vmc_amd.peps.gi.compat import gi_apply

        return gi_apply(self, None, sample)
