"""Gauge removal configuration and API (mirror of src/vmc/gauge/gauge.py:1-37).

The reference declares the configuration and the projection entry point but
does not implement the projection: compute_gauge_projection raises, and so
SRPreconditioner.apply raises when a gauge_config is set. The same holds here,
with the reference's message.
"""
from __future__ import annotations

from dataclasses import dataclass

__all__ = ["GaugeConfig", "compute_gauge_projection"]


@dataclass(frozen=True)
class GaugeConfig:
    """gauge.py:18-23."""

    rcond: float | None = None
    include_global_scale: bool = True


def compute_gauge_projection(cfg: GaugeConfig, model, params, *, return_info: bool = False):
    """gauge.py:26-37: not implemented in the PEPS-only runtime."""
    raise NotImplementedError(
        "Gauge removal is not implemented in the PEPS-only runtime. "
        "Use gauge_config=None."
    )
