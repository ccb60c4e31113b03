"""Utilities (mirror of src/vmc/utils)."""
from vmc_amd.utils.vmc_utils import batched_eval, build_dense_jac, flatten_samples, ravel_params

__all__ = ["batched_eval", "build_dense_jac", "flatten_samples", "ravel_params"]
