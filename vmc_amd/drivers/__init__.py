"""TDVP / imaginary-time driver (mirror of src/vmc/drivers/{integrators,tdvp}.py).

Host loop over the device sampler and SR solve: every derivative evaluation
rebuilds the bottom cache, samples n_samples configurations with
ceil(n_samples / n_chains) sweeps per chain, and solves SR (tdvp.py:113-163).
"""
from __future__ import annotations

import logging
import math

import numpy as np
import torch

from vmc_amd.core import _sample_counts, _trim_samples, make_mc_sampler
from vmc_amd.peps import ChainKeys, build_mc_kernels

logger = logging.getLogger(__name__)

__all__ = ["Euler", "RK4", "RealTimeUnit", "ImaginaryTimeUnit", "TDVPDriver", "SimulationData"]


def _add_scaled(base, delta, scale):
    """utils.py:60-65."""
    return base + scale * delta.to(base.dtype)


class Integrator:
    @staticmethod
    def step(derivative_fn, state, t, dt, carry):
        raise NotImplementedError


class Euler(Integrator):
    """integrators.py:45-63."""

    @staticmethod
    def step(derivative_fn, state, t, dt, carry):
        k, carry_next, aux = derivative_fn(state, t, carry)
        return _add_scaled(state, k, dt), t + dt, carry_next, aux


class RK4(Integrator):
    """integrators.py:66-109 (each stage resamples)."""

    @staticmethod
    def step(derivative_fn, state, t, dt, carry):
        k1, c1, _ = derivative_fn(state, t, carry)
        k2, c2, _ = derivative_fn(_add_scaled(state, k1, 0.5 * dt), t + 0.5 * dt, c1)
        k3, c3, _ = derivative_fn(_add_scaled(state, k2, 0.5 * dt), t + 0.5 * dt, c2)
        k4, c4, aux = derivative_fn(_add_scaled(state, k3, dt), t + dt, c3)
        incr = (k1 + 2.0 * k2 + 2.0 * k3 + k4) / 6.0
        return _add_scaled(state, incr, dt), t + dt, c4, aux


class TimeUnit:
    grad_factor = 1.0

    def default_integrator(self):
        raise NotImplementedError


class RealTimeUnit(TimeUnit):
    """integrators.py:125-133: grad factor -i, RK4."""

    grad_factor = -1.0j

    def default_integrator(self):
        return RK4()


class ImaginaryTimeUnit(TimeUnit):
    """integrators.py:136-144: grad factor -1, Euler."""

    grad_factor = -1.0

    def default_integrator(self):
        return Euler()


class TDVPDriver:
    """tdvp.py:54-249 with the device sampler and SR solve."""

    def __init__(self, model, operator, *, preconditioner, dt, t0=0.0, time_unit=RealTimeUnit(), integrator=None,
                 sampler_seed: int = 0, n_samples: int = 1, n_chains: int = 1, full_gradient: bool = False):
        self._model = model
        self.full_gradient = bool(full_gradient)
        self.operator = operator
        self.preconditioner = preconditioner
        self.dt = float(dt)
        self.t = float(t0)
        self.step_count = 0
        self.n_samples = int(n_samples)
        self.n_chains = int(n_chains)
        self.time_unit = time_unit
        self.integrator = integrator or time_unit.default_integrator()
        self._params = model.flat_params().clone()
        self._key = ChainKeys(int(sampler_seed), 0)
        self._configs = model.random_physical_configuration(int(sampler_seed) + 1, n_samples=n_chains).reshape(
            n_chains, -1).to(model.device)
        # full_gradient: o in the full (d x params) layout and p = None, so the
        # preconditioner solves over the full Jacobian (DirectSolve or QRSolve)
        self._init_cache, transition, estimate = build_mc_kernels(model, operator, full_gradient=self.full_gradient)
        self._mc_sampler = make_mc_sampler(transition, estimate)
        self.residual_error = None
        self._local_energies = None

    def _time_derivative(self, params, t, carry):
        """tdvp.py:113-163."""
        key, configs = carry
        _, _, chain_length, total = _sample_counts(self.n_samples, self.n_chains)
        cache = self._init_cache(params, configs)
        (configs, key, _), (hist, est) = self._mc_sampler(params, configs, key, cache, n_steps=chain_length)
        samples = _trim_samples(hist, total, self.n_samples)
        o = _trim_samples(est.local_log_derivatives, total, self.n_samples)
        e = _trim_samples(est.local_estimate, total, self.n_samples)
        p = (None if est.active_slice_indices is None
             else _trim_samples(est.active_slice_indices, total, self.n_samples))
        updates, metrics = self.preconditioner.apply(self._model, params, samples, o, p, e,
                                                     grad_factor=self.time_unit.grad_factor)
        return updates, (key, configs), (e, metrics)

    def run(self, T: float) -> None:
        """tdvp.py:205-237."""
        duration = float(T)
        assert duration > 0.0, f"T must be positive, got {duration}"
        n_steps = int(round(duration / self.dt))
        eps = 1e-12 * max(1.0, abs(duration), abs(self.dt))
        assert abs(duration - n_steps * self.dt) <= eps, f"T={duration} must be an integer multiple of dt={self.dt}"
        assert n_steps > 0
        carry = (self._key, self._configs)
        for _ in range(n_steps):
            self._params, self.t, carry, (e, metrics) = self.integrator.step(
                self._time_derivative, self._params, self.t, self.dt, carry)
            self._local_energies = e
            self.residual_error = metrics.get("residual_error")
            logger.info("E_mean=%s", float(e.mean().real))
        self._key, self._configs = carry
        self._model.set_flat_params(self._params)
        self.step_count += n_steps

    @property
    def energy(self):
        e = self._local_energies
        if e is None:
            return None
        # nkstats.statistics fields (tdvp.py:226-241); a complex E keeps a complex mean
        m = e.mean()
        var = float((e - m).abs().pow(2).mean()) * e.numel() / max(1, e.numel() - 1)
        mean = complex(m) if e.is_complex() else float(m)
        return {"mean": mean, "variance": var, "error_of_mean": math.sqrt(var / max(1, e.numel()))}

    @property
    def model(self):
        self._model.set_flat_params(self._params)
        return self._model

    # ------------------------------------------------------------------ #
    # Persistence of the driver state: the carry of tdvp.py:217-228
    # (tensors, t, sampler key, sampler configuration) plus step_count.
    # ------------------------------------------------------------------ #

    def state_dict(self) -> dict:
        return {
            "params": self._params.detach().clone(),
            "configs": self._configs.detach().clone(),
            "t": float(self.t),
            "step_count": int(self.step_count),
            "key_seed": int(self._key.seed),
            "key_sweep": int(self._key.sweep),
            "dt": float(self.dt),
        }

    def load_state_dict(self, state: dict) -> None:
        params = state["params"].to(device=self._params.device, dtype=self._params.dtype)
        if params.shape != self._params.shape:
            raise ValueError(f"parameter vector of {params.numel()} entries, the model has {self._params.numel()}")
        configs = state["configs"].to(device=self._configs.device, dtype=self._configs.dtype)
        if configs.shape != self._configs.shape:
            raise ValueError(f"configurations {tuple(configs.shape)}, the driver has {tuple(self._configs.shape)}")
        self._params = params.clone()
        self._configs = configs.clone()
        self.t = float(state["t"])
        self.step_count = int(state["step_count"])
        self._key = ChainKeys(int(state["key_seed"]), int(state["key_sweep"]))
        self._model.set_flat_params(self._params)

    def save(self, path) -> None:
        """Write the state to `path` (safetensors: parameters and configurations)
        and `path`.json (t, step_count, key, dt): files that execute nothing on load."""
        import json
        from pathlib import Path

        from safetensors.torch import save_file

        st = self.state_dict()
        path = Path(path)
        path.parent.mkdir(parents=True, exist_ok=True)
        tensors = {"params": st["params"].cpu().contiguous(), "configs": st["configs"].cpu().contiguous()}
        if tensors["params"].is_complex():
            tensors["params"] = torch.view_as_real(tensors["params"]).contiguous()
        save_file(tensors, str(path))
        meta = {k: st[k] for k in ("t", "step_count", "key_seed", "key_sweep", "dt")}
        meta["params_complex"] = bool(st["params"].is_complex())
        Path(str(path) + ".json").write_text(json.dumps(meta))

    def load(self, path) -> None:
        """Resume from save(path): the next run() continues the sampler stream and
        the parameter trajectory exactly where the saved driver stopped."""
        import json
        from pathlib import Path

        from safetensors.torch import load_file

        tensors = load_file(str(path))
        meta = json.loads(Path(str(path) + ".json").read_text())
        params = tensors["params"]
        if meta.get("params_complex"):
            params = torch.view_as_complex(params.contiguous())
        if abs(float(meta["dt"]) - self.dt) > 1e-15 * max(1.0, abs(self.dt)):
            raise ValueError(f"saved dt {meta['dt']} differs from this driver's {self.dt}")
        self.load_state_dict({"params": params, "configs": tensors["configs"], **meta})


class SimulationData:
    """Observables time series written as JSON (the container of the reference's
    examples, examples/lgt/observables.py:24-101): metadata, then per logged
    step the time, energy statistics and optional observables."""

    _LISTS = ("steps", "times", "energies", "energy_errors", "energy_variances", "electric_energies",
              "plaquette_maps", "vbs_dx", "vbs_dy", "wilson_loops")

    def __init__(self, model_type: str = "", lattice_size=(0, 0), parameters: dict | None = None):
        self.model_type = model_type
        self.lattice_size = tuple(lattice_size)
        self.parameters = dict(parameters or {})
        for k in self._LISTS:
            setattr(self, k, [])

    def add_step(self, step, time, energy, energy_error=0.0, energy_variance=0.0, electric_energy=None):
        self.steps.append(int(step))
        self.times.append(float(time))
        self.energies.append(float(energy))
        self.energy_errors.append(float(energy_error))
        self.energy_variances.append(float(energy_variance))
        if electric_energy is not None:
            self.electric_energies.append(float(electric_energy))

    def add_driver_step(self, driver) -> None:
        """The step, time and energy statistics of a TDVPDriver after run()."""
        e = driver.energy
        mean = e["mean"].real if isinstance(e["mean"], complex) else e["mean"]
        self.add_step(driver.step_count, driver.t, mean, e["error_of_mean"], e["variance"])

    def add_plaquette_map(self, plaq_map):
        self.plaquette_maps.append(np.asarray(plaq_map, dtype=np.float64).tolist())

    def add_vbs_order(self, dx, dy):
        self.vbs_dx.append(float(dx))
        self.vbs_dy.append(float(dy))

    def add_wilson_loops(self, loops):
        self.wilson_loops.append([(int(a), float(v)) for a, v in loops])

    def save(self, path) -> None:
        import json
        from pathlib import Path

        path = Path(path)
        path.parent.mkdir(parents=True, exist_ok=True)
        data = {"model_type": self.model_type, "lattice_size": list(self.lattice_size), "parameters": self.parameters}
        data.update({k: getattr(self, k) for k in self._LISTS})
        path.write_text(json.dumps(data, indent=2))

    @classmethod
    def load(cls, path) -> "SimulationData":
        import json
        from pathlib import Path

        data = json.loads(Path(path).read_text())
        obj = cls(data.get("model_type", ""), tuple(data.get("lattice_size", (0, 0))), data.get("parameters", {}))
        for k in cls._LISTS:
            setattr(obj, k, data.get(k, []))
        return obj
