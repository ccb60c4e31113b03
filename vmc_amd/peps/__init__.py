from vmc_amd.peps.kernels import Cache, ChainKeys, Context, LocalEstimates, build_mc_kernels
from vmc_amd.peps.model import PEPS
from vmc_amd.peps.strategy import ContractionStrategy, DensityMatrix, NoTruncation, Variational, ZipUp

__all__ = ["Cache", "ChainKeys", "Context", "LocalEstimates", "build_mc_kernels", "PEPS", "ContractionStrategy", "DensityMatrix", "NoTruncation", "Variational", "ZipUp"]
