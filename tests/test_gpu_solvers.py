"""DirectSolve solvers, full-Jacobian SR, full-gradient TDVP and driver persistence.

* solve_cg and solve_svd (src/vmc/qgt/solvers.py:16-37) inside
  SRPreconditioner.apply against oracle.sr with the same solver, both spaces;
  the CG step count equals the oracle loop's.
* DirectSolve on a full Jacobian (p = None, preconditioners.py:296-297)
  against oracle.sr.direct_solve_parameter on the full columns.
* TDVPDriver(full_gradient=True) (tdvp.py:70, 146): one step through the full
  Jacobian equals one step through the sliced Jacobian from the same state
  (the expanded sliced Jacobian is the full one), and QRSolve is reachable.
* save / load of the driver state resumes the trajectory bit for bit.
"""
import functools

import numpy as np
import pytest
import torch

from oracle import peps as P
from oracle import sr as S

pytestmark = pytest.mark.gpu


def _synthetic(ns, shape=(3, 3), D=3, seed=11):
    pps = P.params_per_site(shape, D)
    rng = np.random.default_rng(seed)
    o = rng.standard_normal((ns, sum(pps)))
    e = rng.standard_normal(ns)
    p = rng.integers(0, 2, size=o.shape).astype(np.int8)
    return o, p, e, pps


def _model(shape, D, dtype=torch.float64, seed=0):
    from vmc_amd.peps import PEPS, ZipUp

    return PEPS(shape=shape, bond_dim=D, contraction_strategy=ZipUp(D * D), dtype=dtype, seed=seed)


@pytest.mark.parametrize("space", ["parameter", "sample"])
def test_solve_cg_matches_oracle(space):
    from vmc_amd.preconditioners import DirectSolve, SRPreconditioner, solve_cg
    from vmc_amd.qgt import ParameterSpace, SampleSpace

    print(f"[solvers] cg space={space}", flush=True)
    o, p, e, pps = _synthetic(120)
    ref_solver = functools.partial(S.solve_cg, maxiter=400, tol=1e-10)
    ref, _ = S.sr_apply(o, p, e, pps, (2,) * 9, space=space, diag_shift=1e-1, grad_factor=-1.0, solver=ref_solver)
    pre = SRPreconditioner(ParameterSpace() if space == "parameter" else SampleSpace(), diag_shift=1e-1,
                           strategy=DirectSolve(functools.partial(solve_cg, maxiter=400, tol=1e-10)))
    x, m = pre.apply(_model((3, 3), 3), None, None, torch.tensor(o, device="cuda"), torch.tensor(p, device="cuda"),
                     torch.tensor(e, device="cuda"), grad_factor=-1.0)
    err = np.max(np.abs(x.cpu().numpy() - ref)) / np.max(np.abs(ref))
    chol, _ = S.sr_apply(o, p, e, pps, (2,) * 9, space=space, diag_shift=1e-1, grad_factor=-1.0)
    print(f"  cg steps {m['cg_steps']}, rel err vs oracle cg {err:.2e}, "
          f"oracle cg vs cholesky {np.max(np.abs(ref - chol)) / np.max(np.abs(chol)):.2e}", flush=True)
    assert 0 < m["cg_steps"] < 400
    assert err < 1e-8


def test_solve_cg_stops_at_maxiter():
    from vmc_amd.preconditioners import DirectSolve, SRPreconditioner, solve_cg
    from vmc_amd.qgt import SampleSpace

    o, p, e, pps = _synthetic(120)
    ref, _ = S.sr_apply(o, p, e, pps, (2,) * 9, space="sample", diag_shift=1e-3, grad_factor=-1.0,
                        solver=functools.partial(S.solve_cg, maxiter=7, tol=1e-14))
    pre = SRPreconditioner(SampleSpace(), diag_shift=1e-3,
                           strategy=DirectSolve(functools.partial(solve_cg, maxiter=7, tol=1e-14)))
    x, m = pre.apply(_model((3, 3), 3), None, None, torch.tensor(o, device="cuda"), torch.tensor(p, device="cuda"),
                     torch.tensor(e, device="cuda"), grad_factor=-1.0)
    assert m["cg_steps"] == 7
    assert np.max(np.abs(x.cpu().numpy() - ref)) / np.max(np.abs(ref)) < 1e-9


@pytest.mark.parametrize("space", ["parameter", "sample"])
def test_solve_svd_matches_oracle_when_certified(space):
    from vmc_amd._lib import VMCError
    from vmc_amd.preconditioners import DirectSolve, SRPreconditioner, solve_svd
    from vmc_amd.qgt import ParameterSpace, SampleSpace

    print(f"[solvers] svd space={space}", flush=True)
    o, p, e, pps = _synthetic(120)
    sp = ParameterSpace() if space == "parameter" else SampleSpace()
    ref, _ = S.sr_apply(o, p, e, pps, (2,) * 9, space=space, diag_shift=1e-2, grad_factor=-1.0, solver=S.solve_svd)
    pre = SRPreconditioner(sp, diag_shift=1e-2, strategy=DirectSolve(solve_svd))
    x, _ = pre.apply(_model((3, 3), 3), None, None, torch.tensor(o, device="cuda"), torch.tensor(p, device="cuda"),
                     torch.tensor(e, device="cuda"), grad_factor=-1.0)
    err = np.max(np.abs(x.cpu().numpy() - ref)) / np.max(np.abs(ref))
    print(f"  rel err vs oracle svd {err:.2e}", flush=True)
    assert err < 1e-8
    # a cutoff that may bite is refused, not silently ignored
    pre = SRPreconditioner(sp, diag_shift=1e-2, strategy=DirectSolve(functools.partial(solve_svd, rcond=1e-2)))
    with pytest.raises(VMCError, match="solve_svd"):
        pre.apply(_model((3, 3), 3), None, None, torch.tensor(o, device="cuda"), torch.tensor(p, device="cuda"),
                  torch.tensor(e, device="cuda"), grad_factor=-1.0)


def test_full_jacobian_direct_solve_matches_oracle():
    from vmc_amd.preconditioners import SRPreconditioner
    from vmc_amd.qgt import ParameterSpace

    print("[solvers] full-Jacobian DirectSolve", flush=True)
    rng = np.random.default_rng(5)
    ns, n = 80, 150
    O = rng.standard_normal((ns, n))
    e = rng.standard_normal(ns)
    dv = -(e - e.mean()) / ns
    ref, ref_m = S.direct_solve_parameter(O, dv, 1e-2)
    ref = 2.0 * ref.real
    pre = SRPreconditioner(ParameterSpace(), diag_shift=1e-2)
    x, m = pre.apply(None, torch.zeros(n, dtype=torch.float64), None, torch.tensor(O, device="cuda"), None,
                     torch.tensor(e, device="cuda"), grad_factor=-1.0)
    err = np.max(np.abs(x.cpu().numpy() - ref)) / np.max(np.abs(ref))
    print(f"  rel err {err:.2e}", flush=True)
    assert err < 1e-10
    assert abs(m["residual_error"] - ref_m["residual_error"]) < 1e-6 * max(1.0, ref_m["residual_error"]) + 1e-10


def _tfim(shape, h=1.0):
    from vmc_amd.operators import DiagonalTerm, LocalHamiltonian, OneSiteTerm

    zz = -np.array([1.0, -1.0, -1.0, 1.0])
    t = [OneSiteTerm(r, c, -h * np.array([[0.0, 1.0], [1.0, 0.0]])) for r in range(shape[0]) for c in range(shape[1])]
    t += [DiagonalTerm(((r, c), (r, c + 1)), zz) for r in range(shape[0]) for c in range(shape[1] - 1)]
    t += [DiagonalTerm(((r, c), (r + 1, c)), zz) for r in range(shape[0] - 1) for c in range(shape[1])]
    return LocalHamiltonian(shape=shape, terms=tuple(t))


def _driver(full_gradient, strategy=None, seed=2):
    from vmc_amd.drivers import ImaginaryTimeUnit, TDVPDriver
    from vmc_amd.preconditioners import DirectSolve, SRPreconditioner
    from vmc_amd.qgt import ParameterSpace

    shape = (3, 3)
    model = _model(shape, 2, seed=seed)
    pre = SRPreconditioner(ParameterSpace(), diag_shift=1e-2, strategy=strategy or DirectSolve())
    return TDVPDriver(model, _tfim(shape), preconditioner=pre, dt=0.05, time_unit=ImaginaryTimeUnit(),
                      n_samples=96, n_chains=16, sampler_seed=4, full_gradient=full_gradient)


def test_full_gradient_step_equals_sliced_step():
    print("[solvers] full_gradient TDVP step", flush=True)
    a, b = _driver(False), _driver(True)
    a.run(0.05)
    b.run(0.05)
    pa, pb = a._params.cpu().numpy(), b._params.cpu().numpy()
    err = np.max(np.abs(pa - pb)) / np.max(np.abs(pa))
    print(f"  sliced vs full step rel diff {err:.2e}", flush=True)
    assert err < 1e-10
    assert torch.equal(a._configs, b._configs)


def test_full_gradient_qr_solve_runs():
    from vmc_amd.preconditioners import QRSolve

    print("[solvers] full_gradient TDVP with QRSolve", flush=True)
    d = _driver(True, strategy=QRSolve())
    d.run(0.1)
    assert d.step_count == 2 and np.isfinite(d.energy["mean"])
    assert d.preconditioner.residual_error is not None


def test_driver_save_load_resumes_exactly(tmp_path):
    from vmc_amd.drivers import SimulationData

    print("[solvers] driver save/load", flush=True)
    ref = _driver(False)
    ref.run(0.05)
    ref.run(0.05)
    a = _driver(False)
    a.run(0.05)
    a.save(tmp_path / "state.safetensors")
    data = SimulationData("tfim", (3, 3), {"dt": 0.05})
    data.add_driver_step(a)
    data.save(tmp_path / "obs.json")
    b = _driver(False, seed=9)  # other initial tensors: everything must come from the file
    b.load(tmp_path / "state.safetensors")
    assert b.step_count == 1 and abs(b.t - 0.05) < 1e-15
    b.run(0.05)
    assert torch.equal(b._params, ref._params)
    assert torch.equal(b._configs, ref._configs)
    assert b.step_count == ref.step_count == 2
    back = SimulationData.load(tmp_path / "obs.json")
    assert back.steps == [1] and back.energies == data.energies
