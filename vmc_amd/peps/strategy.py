"""Boundary-MPS contraction strategies (selectors for the device engine).

Same class names and constructor arguments as src/vmc/peps/common/strategy.py:25-92.
On the device the strategy picks the truncation performed by the zip-up kernels:
* ``ZipUp(chi)``: keep the top-chi left singular subspace of each theta matrix;
* ``DensityMatrix(chi)``: keep the top-min(chi, n) eigenvectors V of the n x n
  theta^H theta, new tensor theta V and carry V^H (strategy.py:153-176). The
  singular values stay on the new tensor instead of the carry, so every cut
  after the first truncates a different matrix than zip-up does. When
  rank(theta) < k < n the kept vectors include arbitrary null-space vectors of
  theta; a later truncating cut then depends on that choice, in the reference
  as here, so amplitudes are only reproducible where no such cut occurs;
* ``NoTruncation()``: zip-up keeping every singular vector (exact contraction);
* ``Variational(chi, n_sweeps=2)``: QR-initialised one-site ALS compression
  (strategy.py:179-427), the reference's default strategy.
"""
from __future__ import annotations

from vmc_amd import _lib


class ContractionStrategy:
    def __init__(self, truncate_bond_dimension: int):
        self.truncate_bond_dimension = int(truncate_bond_dimension)

    device_code = None

    def __repr__(self):
        return f"{type(self).__name__}({self.truncate_bond_dimension})"


class NoTruncation(ContractionStrategy):
    device_code = _lib.VMC_NO_TRUNCATION

    def __init__(self):
        super().__init__(truncate_bond_dimension=1)


class ZipUp(ContractionStrategy):
    device_code = _lib.VMC_ZIPUP


class DensityMatrix(ContractionStrategy):
    device_code = _lib.VMC_DENSITY_MATRIX


class Variational(ContractionStrategy):
    device_code = _lib.VMC_VARIATIONAL

    def __init__(self, truncate_bond_dimension: int, n_sweeps: int = 2):
        super().__init__(truncate_bond_dimension)
        self.n_sweeps = int(n_sweeps)


def device_strategy(strategy):
    if strategy.device_code is None:
        raise NotImplementedError(f"{strategy!r} has no gfx950 kernel; use ZipUp or DensityMatrix")
    return strategy.device_code
