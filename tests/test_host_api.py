"""Host-side API pieces that need no device: solver selection of DirectSolve,
the gauge API (a stub in the reference, gauge/gauge.py:26-37) and the
observables JSON container (examples/lgt/observables.py:24-101)."""
import functools

import pytest


def test_direct_solve_solver_resolution():
    from vmc_amd import _lib
    from vmc_amd.preconditioners import DirectSolve, _solver_args, solve_cg, solve_cholesky, solve_svd

    assert DirectSolve().solver is solve_cholesky
    assert _solver_args(solve_cholesky) == (_lib.VMC_SOLVE_CHOLESKY, 0.0, 0)
    assert _solver_args(solve_cg) == (_lib.VMC_SOLVE_CG, 1e-5, 1000)  # solvers.py:16-18 defaults
    assert _solver_args(functools.partial(solve_cg, maxiter=50, tol=1e-9)) == (_lib.VMC_SOLVE_CG, 1e-9, 50)
    assert _solver_args(solve_svd) == (_lib.VMC_SOLVE_SVD, 1e-12, 0)  # solvers.py:30 default rcond
    assert _solver_args(functools.partial(solve_svd, rcond=1e-6)) == (_lib.VMC_SOLVE_SVD, 1e-6, 0)
    with pytest.raises(NotImplementedError):
        _solver_args(lambda m, r: r)
    with pytest.raises(NotImplementedError):
        solve_cg(None, None)


def test_gauge_projection_is_the_reference_stub():
    from vmc_amd.gauge import GaugeConfig, compute_gauge_projection
    from vmc_amd.preconditioners import SRPreconditioner

    cfg = GaugeConfig()
    assert cfg.rcond is None and cfg.include_global_scale is True
    with pytest.raises(NotImplementedError, match="Use gauge_config=None"):
        compute_gauge_projection(cfg, None, {})
    pre = SRPreconditioner(gauge_config=cfg)  # accepted at construction, as in the reference
    with pytest.raises(NotImplementedError, match="Gauge removal"):
        pre.apply(None, None, None, None, None, None)


def test_simulation_data_round_trip(tmp_path):
    from vmc_amd.drivers import SimulationData

    d = SimulationData("z2_pure_gauge", (4, 4), {"g": 0.5})
    d.add_step(1, 0.01, -3.5, 0.01, 0.2, electric_energy=-1.25)
    d.add_step(2, 0.02, -3.6)
    d.add_plaquette_map([[0.1, 0.2], [0.3, 0.4]])
    d.add_vbs_order(0.01, -0.02)
    d.add_wilson_loops([(1, 0.9), (4, 0.5)])
    d.save(tmp_path / "o.json")
    b = SimulationData.load(tmp_path / "o.json")
    assert b.model_type == "z2_pure_gauge" and tuple(b.lattice_size) == (4, 4) and b.parameters == {"g": 0.5}
    assert b.steps == [1, 2] and b.energies == [-3.5, -3.6] and b.electric_energies == [-1.25]
    assert b.plaquette_maps == [[[0.1, 0.2], [0.3, 0.4]]] and b.vbs_dx == [0.01]
    assert [tuple(x) for x in b.wilson_loops[0]] == [(1, 0.9), (4, 0.5)]
