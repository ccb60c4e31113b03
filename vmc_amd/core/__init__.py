from vmc_amd.core.sampling import _sample_counts, _trim_samples, make_mc_sampler

__all__ = ["make_mc_sampler", "_sample_counts", "_trim_samples"]
