"""Exact-enumeration imaginary-time evolution to the ground state, end to end.

The recipe of the reference's physics tests (tests/test_physics_models.py:32-99,
246-283): every basis state is a sample, weighted by |psi|^2; SR with a
diagonal shift gives the update (the test's own NumPy solve, as there), and
the tensors step in imaginary time. The device supplies the amplitudes, the
full dpsi/dA and E_loc; the same loop on the oracle supplies the reference
trajectory. In FP64 the two energy trajectories must agree to 1e-8 at every
step, and the 2x3 Heisenberg energy must reach the exact ground-state energy
(dense H) to 1e-3.
"""
import numpy as np
import pytest
import torch

from oracle import exact as E
from oracle import hamiltonians as HH
from oracle import peps as P

pytestmark = pytest.mark.gpu


def _sr_step(amp, grad, eloc, shift, dt, flat):
    w = np.abs(amp) ** 2
    w /= w.sum()
    o = grad / amp[:, None]
    om = (w[:, None] * o).sum(0)
    wo = o * np.sqrt(w[:, None])
    S = wo.conj().T @ wo - np.outer(om.conj(), om)
    em = (w * eloc).sum()
    f = o.conj().T @ (w * (eloc - em))
    x = np.linalg.solve(S + shift * np.eye(len(S)), -f)
    return flat + dt * x.real, float(np.real(em))


def _unflat(flat, T):
    out, off = [], 0
    for row in T:
        r = []
        for t in row:
            r.append(flat[off:off + t.size].reshape(t.shape))
            off += t.size
        out.append(r)
    return out


def test_imaginary_time_reaches_the_heisenberg_ground_state():
    from vmc_amd.operators import HorizontalTwoSiteTerm, LocalHamiltonian, VerticalTwoSiteTerm
    from vmc_amd.peps import PEPS, NoTruncation
    from vmc_amd.peps.compat import _value_and_grad, local_estimate

    shape, D, steps, dt, shift = (2, 3), 3, 40, 0.1, 1e-3
    print(f"[imag] exact imaginary time {shape} D={D} Heisenberg, {steps} steps", flush=True)
    terms = HH.heisenberg(shape)
    e0 = float(np.linalg.eigvalsh(E.dense_hamiltonian(terms, shape)).min())
    dev_terms = [HorizontalTwoSiteTerm(r, c, HH.HEISENBERG_BOND) for r in range(shape[0]) for c in range(shape[1] - 1)]
    dev_terms += [VerticalTwoSiteTerm(r, c, HH.HEISENBERG_BOND) for r in range(shape[0] - 1) for c in range(shape[1])]
    H = LocalHamiltonian(shape=shape, terms=tuple(dev_terms))
    T = P.random_peps(shape, D, seed=0)
    model = PEPS(shape=shape, bond_dim=D, contraction_strategy=NoTruncation(), tensors=T, dtype=torch.float64)
    states = E.all_configs(shape[0] * shape[1])
    flat_dev = np.concatenate([t.ravel() for row in T for t in row])
    flat_ref = flat_dev.copy()
    s_dev = torch.tensor(states)
    for it in range(steps):
        model.set_flat_params(torch.tensor(flat_dev, device="cuda"))
        amp, grad, _ = _value_and_grad(model, s_dev, full_gradient=True)
        amp, grad = amp.cpu().numpy(), grad.cpu().numpy()
        keep = np.abs(amp) > 1e-12
        eloc = local_estimate(model, s_dev[torch.tensor(keep)], H).cpu().numpy()
        flat_dev, e_dev = _sr_step(amp[keep], grad[keep], eloc, shift, dt, flat_dev)

        TT = _unflat(flat_ref, T)
        ra, rg, _ = P.value_and_grad(TT, states, shape, P.NoTruncation(), D, full_gradient=True)
        rk = np.abs(ra) > 1e-12
        re = P.local_estimate(TT, states[rk], shape, P.NoTruncation(), terms, ra[rk])
        flat_ref, e_ref = _sr_step(ra[rk], rg[rk], re, shift, dt, flat_ref)
        if it % 10 == 0 or it == steps - 1:
            print(f"  step {it}: E device {e_dev:.10f} oracle {e_ref:.10f} (E0 {e0:.6f})", flush=True)
        assert abs(e_dev - e_ref) <= 1e-8 * max(1.0, abs(e_ref))
    assert abs(e_dev - e0) <= 1e-3 * abs(e0)


def test_real_time_rk4_conserves_the_energy():
    """Real-time TDVP (grad factor -i, RK4) on a complex PEPS, exact enumeration:
    the device trajectory equals the oracle's, and the energy is conserved to the
    order of the time-step and regularisation errors (4e-6 over 20 steps, against
    an energy scale of 16)."""
    from vmc_amd.operators import DiagonalTerm, LocalHamiltonian, OneSiteTerm
    from vmc_amd.peps import PEPS, NoTruncation
    from vmc_amd.peps.compat import _value_and_grad, local_estimate

    shape, D, steps, dt, shift, J, h = (2, 3), 2, 20, 0.02, 1e-4, 1.0, 1.5
    print(f"[real] exact real-time RK4 {shape} D={D} complex TFIM, {steps} steps", flush=True)
    terms = HH.tfim(shape, J, h)
    zz = -J * np.array([1.0, -1.0, -1.0, 1.0])
    ox = -h * np.array([[0.0, 1.0], [1.0, 0.0]])
    dev_terms = [OneSiteTerm(r, c, ox) for r in range(shape[0]) for c in range(shape[1])]
    dev_terms += [DiagonalTerm(((r, c), (r, c + 1)), zz) for r in range(shape[0]) for c in range(shape[1] - 1)]
    dev_terms += [DiagonalTerm(((r, c), (r + 1, c)), zz) for r in range(shape[0] - 1) for c in range(shape[1])]
    H = LocalHamiltonian(shape=shape, terms=tuple(dev_terms))
    T = P.random_peps(shape, D, seed=0, dtype=np.complex128)
    model = PEPS(shape=shape, bond_dim=D, contraction_strategy=NoTruncation(), tensors=T, dtype=torch.complex128)
    states = E.all_configs(shape[0] * shape[1])
    s_dev = torch.tensor(states)

    def sr(amp, grad, eloc):
        w = np.abs(amp) ** 2
        w /= w.sum()
        o = grad / amp[:, None]
        om = (w[:, None] * o).sum(0)
        wo = o * np.sqrt(w[:, None])
        S = wo.conj().T @ wo - np.outer(om.conj(), om)
        em = (w * eloc).sum()
        f = o.conj().T @ (w * (eloc - em))
        return np.linalg.solve(S + shift * np.eye(len(S)), -1j * f), float(np.real(em))

    def f_dev(x):
        model.set_flat_params(torch.tensor(x, device="cuda"))
        amp, grad, _ = _value_and_grad(model, s_dev, full_gradient=True)
        amp, grad = amp.cpu().numpy(), grad.cpu().numpy()
        keep = np.abs(amp) > 1e-12
        eloc = local_estimate(model, s_dev[torch.tensor(keep)], H).cpu().numpy()
        return sr(amp[keep], grad[keep], eloc)

    def f_ref(x):
        TT = _unflat(x, T)
        amp, grad, _ = P.value_and_grad(TT, states, shape, P.NoTruncation(), D, full_gradient=True)
        keep = np.abs(amp) > 1e-12
        eloc = P.local_estimate(TT, states[keep], shape, P.NoTruncation(), terms, amp[keep])
        return sr(amp[keep], grad[keep], eloc)

    def rk4(f, x):
        k1, e = f(x)
        k2, _ = f(x + 0.5 * dt * k1)
        k3, _ = f(x + 0.5 * dt * k2)
        k4, _ = f(x + dt * k3)
        return x + dt / 6 * (k1 + 2 * k2 + 2 * k3 + k4), e

    x_dev = np.concatenate([t.ravel() for row in T for t in row])
    x_ref = x_dev.copy()
    energies = []
    for it in range(steps):
        x_dev, e_dev = rk4(f_dev, x_dev)
        x_ref, e_ref = rk4(f_ref, x_ref)
        energies.append(e_dev)
        if it % 5 == 0 or it == steps - 1:
            print(f"  step {it}: E device {e_dev:.12f} oracle {e_ref:.12f}", flush=True)
        assert abs(e_dev - e_ref) <= 1e-9 * max(1.0, abs(e_ref))
    # |H| is about J * (7 bonds) + h * (6 sites) = 16 here; the drift is measured against that scale
    assert np.max(np.abs(np.asarray(energies) - energies[0])) <= 1e-6 * 16.0
