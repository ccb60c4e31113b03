"""MI355X-native (gfx950) PEPS amplitude engine.

Python host code over the C ABI of ``libvmc_amd.so`` (include/vmc_amd.h). The
public names mirror the reference's evaluator, sampler and driver API
(fliingelephant/VMC, src/vmc).
"""
__all__ = []
