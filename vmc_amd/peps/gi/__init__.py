"""Gauge-invariant PEPS (mirror of src/vmc/peps/gi)."""
from vmc_amd.peps.gi.compat import gi_apply, gi_value_and_grad
from vmc_amd.peps.gi.kernels import build_mc_kernels
from vmc_amd.peps.gi.local_terms import (
    GILocalHamiltonian,
    LinkDiagonalTerm,
    MatterMassTerm,
    build_electric_terms,
)
from vmc_amd.peps.gi.model import GIPEPS, GIPEPSConfig

__all__ = ["GIPEPS", "GIPEPSConfig", "GILocalHamiltonian", "LinkDiagonalTerm", "MatterMassTerm",
           "build_electric_terms", "build_mc_kernels", "gi_apply", "gi_value_and_grad"]
