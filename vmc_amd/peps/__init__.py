from vmc_amd.peps.model import PEPS
from vmc_amd.peps.strategy import ContractionStrategy, DensityMatrix, NoTruncation, Variational, ZipUp

__all__ = ["PEPS", "ContractionStrategy", "DensityMatrix", "NoTruncation", "Variational", "ZipUp"]
