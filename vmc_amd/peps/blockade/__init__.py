"""Blockade PEPS (mirror of src/vmc/peps/blockade)."""
from vmc_amd.peps.blockade.compat import blockade_apply, blockade_value_and_grad
from vmc_amd.peps.blockade.hamiltonian import rydberg_hamiltonian
from vmc_amd.peps.blockade.kernels import build_mc_kernels
from vmc_amd.peps.blockade.model import BlockadePEPS, BlockadePEPSConfig, random_independent_set

__all__ = ["BlockadePEPS", "BlockadePEPSConfig", "random_independent_set", "rydberg_hamiltonian", "blockade_apply",
           "blockade_value_and_grad", "build_mc_kernels"]
