"""GPU parity of the SR/QGT layer against oracle.sr (FP64).

vmc_sr_solve vs SRPreconditioner.apply restated (preconditioners.py:268-330) in
both spaces, vmc_qgt_dense vs QGT.to_dense (qgt.py:173-223), and the TDVP
driver plumbing (recipe of tests/test_tdvp_driver.py:38-91).
"""
import numpy as np
import pytest
import torch

from oracle import peps as P
from oracle import sr as S

pytestmark = pytest.mark.gpu


def _data(shape=(3, 3), D=2, ns=96, seed=0):
    T = P.random_peps(shape, D, seed=seed)
    cfg = P.random_configs(shape[0] * shape[1], ns, seed=seed)
    amp, grad, p = P.value_and_grad(T, cfg, shape, P.NoTruncation(), D)
    o = grad / amp[:, None]
    e = np.random.default_rng(seed).standard_normal(ns)
    pps = P.params_per_site(shape, D)
    return o, p, e, pps


def _model(shape, D):
    from vmc_amd.peps import PEPS, ZipUp

    return PEPS(shape=shape, bond_dim=D, contraction_strategy=ZipUp(D * D), dtype=torch.float64)


@pytest.mark.parametrize("space", ["parameter", "sample"])
def test_sr_solve_matches_oracle(space):
    from vmc_amd.preconditioners import SRPreconditioner
    from vmc_amd.qgt import ParameterSpace, SampleSpace

    print(f"[sr] space={space}", flush=True)
    shape, D = (3, 3), 2
    o, p, e, pps = _data(shape, D)
    ref, ref_m = S.sr_apply(o, p, e, pps, (2,) * 9, space=space, diag_shift=1e-3, grad_factor=-1.0)
    pre = SRPreconditioner(ParameterSpace() if space == "parameter" else SampleSpace(), diag_shift=1e-3)
    x, m = pre.apply(_model(shape, D), None, None, torch.tensor(o, device="cuda"),
                     torch.tensor(p, device="cuda"), torch.tensor(e, device="cuda"), grad_factor=-1.0)
    x = x.cpu().numpy()
    err = np.max(np.abs(x - ref)) / np.max(np.abs(ref))
    print(f"  rel err {err:.2e} residual {m['residual_error']:.2e} (oracle {ref_m['residual_error']:.2e})", flush=True)
    # S is singular below the shift (rank <= Ns-1 < d*Np here), so the shifted
    # solve amplifies rounding by ~ lambda_max / diag_shift in parameter space.
    assert err < (1e-7 if space == "parameter" else 1e-9)
    assert abs(m["residual_error"] - ref_m["residual_error"]) < 1e-6 * max(1.0, ref_m["residual_error"]) + 1e-10


def test_qgt_dense_matches_oracle():
    from vmc_amd.qgt import QGT, ParameterSpace, SampleSpace, SlicedJacobian

    print("[sr] qgt dense", flush=True)
    o, p, e, pps = _data()
    jac = SlicedJacobian(torch.tensor(o, device="cuda"), torch.tensor(p, device="cuda"), (2,) * 9,
                         params_per_site=tuple(pps))
    O = S.dense_blocks(o, p, (2,) * 9, S.SiteOrdering(tuple(pps)))
    Sd = QGT(jac, ParameterSpace()).to_dense().cpu().numpy()
    Td = QGT(jac, SampleSpace()).to_dense().cpu().numpy()
    assert np.allclose(Sd, S.qgt_dense_parameter(O), atol=1e-12 * np.abs(Sd).max())
    assert np.allclose(Td, S.qgt_dense_sample(O), atol=1e-12 * np.abs(Td).max())


def test_tdvp_driver_runs_and_counts_steps():
    from oracle.hamiltonians import tfim
    from vmc_amd.drivers import ImaginaryTimeUnit, TDVPDriver
    from vmc_amd.operators import DiagonalTerm, LocalHamiltonian, OneSiteTerm
    from vmc_amd.preconditioners import SRPreconditioner
    from vmc_amd.qgt import SampleSpace

    print("[sr] tdvp driver", flush=True)
    shape = (3, 3)
    terms = []
    for r in range(3):
        for c in range(3):
            terms.append(OneSiteTerm(r, c, -1.0 * np.array([[0.0, 1.0], [1.0, 0.0]])))
            if c + 1 < 3:
                terms.append(DiagonalTerm(((r, c), (r, c + 1)), -np.array([1.0, -1, -1, 1])))
            if r + 1 < 3:
                terms.append(DiagonalTerm(((r, c), (r + 1, c)), -np.array([1.0, -1, -1, 1])))
    op = LocalHamiltonian(shape=shape, terms=tuple(terms))
    model = _model(shape, 2)
    drv = TDVPDriver(model, op, preconditioner=SRPreconditioner(SampleSpace(), diag_shift=1e-2), dt=0.05,
                     time_unit=ImaginaryTimeUnit(), n_samples=128, n_chains=16, sampler_seed=3)
    drv.run(0.1)
    e1 = drv.energy["mean"]
    drv.run(0.15)
    assert drv.step_count == 5 and abs(drv.t - 0.25) < 1e-12
    for _ in range(6):
        drv.run(0.05)
    e2 = drv.energy["mean"]
    print(f"  E after 2 steps {e1:.4f}, after 11 steps {e2:.4f}", flush=True)
    assert np.isfinite(e2) and e2 < e1
