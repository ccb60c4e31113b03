"""Operator terms on the device: complex operator values and per-caller tables.

* Complex, non-Hermitian one-site, horizontal, vertical and diagonal terms on a
  complex PEPS (vmc_plan_set_terms_c; the reference's terms take any dtype,
  src/vmc/operators/local_terms.py:22-200): E_loc against the oracle, whose
  complex E_loc is pinned to the dense Hamiltonian in tests/test_oracle_energy.py.
* A kernel bundle owns its operator tables and its transition's top
  boundaries (Context): a local_estimate with another operator, and a second
  build_mc_kernels with another operator, between transition and estimate
  leave that bundle's o and E_loc unchanged.
"""
import numpy as np
import pytest
import torch

from oracle import hamiltonians as HH
from oracle import peps as P

pytestmark = pytest.mark.gpu


def _complex_terms(shape, seed):
    from vmc_amd.operators import (DiagonalTerm, HorizontalTwoSiteTerm, LocalHamiltonian, OneSiteTerm,
                                   VerticalTwoSiteTerm)

    rng = np.random.default_rng(seed)
    c = lambda *s: rng.standard_normal(s) + 1j * rng.standard_normal(s)
    R, C = shape
    dev, ora = [], P.BucketedTerms(shape)
    for r in range(R):
        for col in range(C):
            o1 = c(2, 2)
            dev.append(OneSiteTerm(r, col, o1))
            ora.one_site[r][col].append(o1)
            if col + 1 < C:
                o2 = c(4, 4)
                dev.append(HorizontalTwoSiteTerm(r, col, o2))
                ora.horizontal[r][col].append(o2)
            if r + 1 < R:
                o3 = c(4, 4)
                dev.append(VerticalTwoSiteTerm(r, col, o3))
                ora.vertical[r][col].append(o3)
                dg = c(4)
                dev.append(DiagonalTerm(((r, col), (r + 1, col)), dg))
                ora.diagonal.append((((r, col), (r + 1, col)), dg))
    return LocalHamiltonian(shape=shape, terms=tuple(dev)), ora


@pytest.mark.parametrize("dtype,tol", [(torch.complex128, 1e-9), (torch.complex64, 1e-4)])
def test_complex_operator_eloc(dtype, tol):
    from vmc_amd.peps import PEPS, ZipUp
    from vmc_amd.peps.compat import local_estimate

    shape, D, chi = (4, 4), 2, 4
    print(f"[terms] complex operators {shape} D={D} chi={chi} {dtype}", flush=True)
    T = P.random_peps(shape, D, seed=31, dtype=np.complex128)
    model = PEPS(shape=shape, bond_dim=D, contraction_strategy=ZipUp(chi), tensors=T, dtype=dtype)
    op, ora = _complex_terms(shape, 32)
    cfg = P.random_configs(16, 64, seed=33)
    e = local_estimate(model, torch.tensor(cfg), op).cpu().numpy()
    ra = P.value(T, cfg, shape, P.ZipUp(chi))
    re = P.local_estimate(T, cfg, shape, P.ZipUp(chi), ora, ra)
    err = np.max(np.abs(e - re)) / max(1.0, np.abs(re).max())
    print(f"  max rel E_loc err {err:.2e}", flush=True)
    assert np.max(np.abs(re.imag)) > 1e-3
    assert err < tol


def test_complex_operator_on_real_plan_is_refused():
    from vmc_amd._lib import VMCError
    from vmc_amd.peps import PEPS, ZipUp
    from vmc_amd.peps.compat import local_estimate

    model = PEPS(shape=(2, 2), bond_dim=2, contraction_strategy=ZipUp(4), dtype=torch.float64, seed=1)
    op, _ = _complex_terms((2, 2), 3)
    with pytest.raises(VMCError, match="complex plan"):
        local_estimate(model, torch.zeros((1, 4), dtype=torch.int32), op)


def _tfim(shape, h):
    from vmc_amd.operators import DiagonalTerm, LocalHamiltonian, OneSiteTerm

    zz = -np.array([1.0, -1.0, -1.0, 1.0])
    ox = -h * np.array([[0.0, 1.0], [1.0, 0.0]])
    R, C = shape
    t = [OneSiteTerm(r, c, ox) for r in range(R) for c in range(C)]
    t += [DiagonalTerm(((r, c), (r, c + 1)), zz) for r in range(R) for c in range(C - 1)]
    t += [DiagonalTerm(((r, c), (r + 1, c)), zz) for r in range(R - 1) for c in range(C)]
    return LocalHamiltonian(shape=shape, terms=tuple(t))


def _heis(shape):
    from vmc_amd.operators import HorizontalTwoSiteTerm, LocalHamiltonian, VerticalTwoSiteTerm

    R, C = shape
    t = [HorizontalTwoSiteTerm(r, c, HH.HEISENBERG_BOND) for r in range(R) for c in range(C - 1)]
    t += [VerticalTwoSiteTerm(r, c, HH.HEISENBERG_BOND) for r in range(R - 1) for c in range(C)]
    return LocalHamiltonian(shape=shape, terms=tuple(t))


def test_bundles_do_not_share_terms_or_top_boundaries():
    from vmc_amd.peps import PEPS, ChainKeys, ZipUp, build_mc_kernels
    from vmc_amd.peps.compat import local_estimate

    shape, D, chi, n = (4, 4), 2, 4, 16
    print(f"[terms] bundle isolation {shape}", flush=True)
    T = P.random_peps(shape, D, seed=41)
    model = PEPS(shape=shape, bond_dim=D, contraction_strategy=ZipUp(chi), tensors=T, dtype=torch.float64)
    cfg0 = torch.tensor(P.random_configs(16, n, seed=42))

    def run(interfere):
        init_cache, transition, estimate = build_mc_kernels(model, _heis(shape))
        cache = init_cache(None, cfg0)
        cfg1, key, ctx = transition(None, cfg0, ChainKeys(5), cache)
        if interfere:
            # another operator through local_estimate and a second bundle, and a
            # second transition of other chains on that bundle
            local_estimate(model, torch.tensor(P.random_configs(16, 3 * n, seed=43)), _tfim(shape, 2.0))
            i2, t2, e2 = build_mc_kernels(model, _tfim(shape, 3.0))
            c2 = i2(None, torch.tensor(P.random_configs(16, 2 * n, seed=44)))
            t2(None, torch.tensor(P.random_configs(16, 2 * n, seed=44)), ChainKeys(9), c2)
            # a second transition on the same bundle with other chains before the first estimate
            c3 = init_cache(None, torch.tensor(P.random_configs(16, n, seed=45)))
            transition(None, torch.tensor(P.random_configs(16, n, seed=45)), ChainKeys(11), c3)
        _, est = estimate(None, cfg1, ctx)
        torch.cuda.synchronize()
        return cfg1.cpu(), est.local_log_derivatives.cpu(), est.local_estimate.cpu()

    a = run(False)
    b = run(True)
    assert torch.equal(a[0], b[0])
    assert torch.equal(a[1], b[1])
    assert torch.equal(a[2], b[2])
    # and both equal the oracle's sweep
    cache_ref = P.init_cache(T, cfg0.numpy(), shape, P.ZipUp(chi))
    _, _, h = P.mc_sampler(T, cfg0.numpy(), cache_ref, shape, P.ZipUp(chi), HH.heisenberg(shape), 5, 1)
    assert np.array_equal(a[0].numpy(), h["samples"][0])
    assert np.allclose(a[2].numpy(), h["eloc"][0], rtol=1e-9, atol=1e-10)
