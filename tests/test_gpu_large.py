"""Truncations beyond the register-resident kernels (trunc_g.hip) and the
configurations that need them, against the FP64 oracle (LAPACK SVD zip-up).

* real theta with more than 32 kept vectors (6x6 D=4 chi=40: 160 x 160 cuts);
* complex theta above 64 rows (6x6 D=3 chi=24: 72 x 72 cuts);
* the real-time TDVP configuration's shape, 12x12 OBC D=6 chi=36 with complex
  216 x 216 cuts, in complex128 and complex64 on a small batch;
* the precision-form sampler (8x8 D=4 chi=16 complex, TFIM h=3.04): identical
  sample sequences in complex128, E_loc error of complex64.
"""
import numpy as np
import pytest
import torch

from oracle import hamiltonians as HH
from oracle import peps as P

pytestmark = pytest.mark.gpu


def _flat(T):
    return np.concatenate([t.ravel() for row in T for t in row])


def _parity(shape, D, chi, dtype, B, seed, tol_log, tol_o, grad=True):
    from vmc_amd.engine import AmpGradPlan

    cplx = dtype.is_complex
    print(f"[large] {shape} D={D} chi={chi} {dtype} B={B}", flush=True)
    T = P.random_peps(shape, D, seed=seed, dtype=np.complex128 if cplx else np.float64)
    cfg = P.random_configs(shape[0] * shape[1], B, seed=seed)
    plan = AmpGradPlan(shape, D, chi, dtype=dtype)
    la, o, _ = plan.amp_grad(torch.tensor(_flat(T), dtype=dtype, device="cuda"), torch.tensor(cfg), grad=grad)
    la = la.cpu().numpy()
    if grad:
        amp, g, _ = P.value_and_grad(T, cfg, shape, P.ZipUp(chi), D)
    else:
        amp = P.value(T, cfg, shape, P.ZipUp(chi))
    ref = np.log(np.abs(amp))
    err_log = np.max(np.abs(la[:, 0] - ref) / np.maximum(1.0, np.abs(ref)))
    dph = np.abs(np.angle(np.exp(1j * (la[:, 1] - np.angle(amp)))))
    msg = f"  rel err log|psi| {err_log:.2e} phase {dph.max():.2e}"
    if grad:
        o_ref = g / amp[:, None]
        err_o = np.max(np.abs(o.cpu().numpy() - o_ref)) / np.max(np.abs(o_ref))
        msg += f" o {err_o:.2e}"
        assert err_o < tol_o
    print(msg, flush=True)
    assert err_log < tol_log
    assert dph.max() < 1e3 * tol_log


def test_real_chi40_6x6():
    _parity((6, 6), 4, 40, torch.float64, 6, 51, 1e-11, 1e-8)


def test_complex_72_rows_6x6():
    _parity((6, 6), 3, 24, torch.complex128, 6, 52, 1e-11, 1e-8)


def test_tdvp_shape_12x12_D6_chi36_c128():
    _parity((12, 12), 6, 36, torch.complex128, 2, 53, 1e-10, 1e-7)


def test_tdvp_shape_12x12_D6_chi36_c64_vs_conditioning():
    """complex64 contractions against the complex128 path (pinned above) on the
    same complex64-representable inputs. As at the 10x10 metric shape, the
    difference is the conditioning of the truncated amplitude: the complex128
    path alone moves log psi by the same order when fed the unrounded draw."""
    from vmc_amd.engine import AmpGradPlan

    shape, D, chi, B = (12, 12), 6, 36, 16
    print(f"[large] {shape} D={D} chi={chi} complex64 vs complex128 B={B}", flush=True)
    T = P.random_peps(shape, D, seed=53, dtype=np.complex128)
    draw = _flat(T)
    cast = draw.astype(np.complex64).astype(np.complex128)
    cfg = torch.tensor(P.random_configs(144, B, seed=54), device="cuda")
    la = {}
    for name, dt, vals in (("c64", torch.complex64, cast), ("c128", torch.complex128, cast),
                           ("c128_draw", torch.complex128, draw)):
        plan = AmpGradPlan(shape, D, chi, dtype=dt)
        la[name] = plan.amp_grad(torch.tensor(vals, dtype=dt, device="cuda"), cfg, grad=False)[0].cpu().numpy()
    ref = np.maximum(1.0, np.abs(la["c128"][:, 0]))
    e32 = np.abs(la["c64"][:, 0] - la["c128"][:, 0]) / ref
    ein = np.abs(la["c128_draw"][:, 0] - la["c128"][:, 0]) / ref
    print(f"  complex64 error median/max {np.median(e32):.2e} {e32.max():.2e}; "
          f"input-rounding change median/max {np.median(ein):.2e} {ein.max():.2e}", flush=True)
    assert e32.max() <= 10 * ein.max() + 1e-6


@pytest.mark.parametrize("dtype", [torch.complex128, torch.complex64])
def test_precision_form_sampler_8x8(dtype):
    from vmc_amd.engine import AmpGradPlan
    from vmc_amd.operators import DiagonalTerm, LocalHamiltonian, OneSiteTerm
    from vmc_amd.operators.local_terms import to_tables

    shape, D, chi, n_chains, seed = (8, 8), 4, 16, 8, 11
    print(f"[large] precision-form sampler 8x8 D4 chi16 TFIM {dtype}", flush=True)
    T = P.random_peps(shape, D, seed=seed, dtype=np.complex128)
    cfg0 = P.random_configs(64, n_chains, seed=seed + 1)
    zz = -np.array([1.0, -1.0, -1.0, 1.0])
    ox = -3.04 * np.array([[0.0, 1.0], [1.0, 0.0]])
    terms = [OneSiteTerm(r, c, ox) for r in range(8) for c in range(8)]
    terms += [DiagonalTerm(((r, c), (r, c + 1)), zz) for r in range(8) for c in range(7)]
    terms += [DiagonalTerm(((r, c), (r + 1, c)), zz) for r in range(7) for c in range(8)]
    plan = AmpGradPlan(shape, D, chi, dtype=dtype)
    plan.set_terms(to_tables(LocalHamiltonian(shape=shape, terms=tuple(terms)), 2))
    params = torch.tensor(_flat(T), dtype=dtype, device="cuda")
    cfg = torch.tensor(cfg0, dtype=torch.int32, device="cuda")
    cache = plan.new_cache(n_chains)
    plan.init_cache(params, cfg, cache)
    _, _, _, e = plan.mc_sweep(params, cfg, cache, seed, 0)
    cref = P.init_cache(T, cfg0, shape, P.ZipUp(chi))
    _, _, hist = P.mc_sampler(T, cfg0, cref, shape, P.ZipUp(chi), HH.tfim(shape, 1.0, 3.04), seed, 1)
    e = e.cpu().numpy()
    if dtype == torch.complex128:
        assert np.array_equal(cfg.cpu().numpy(), hist["samples"][0])
        err = np.max(np.abs(e - hist["eloc"][0])) / max(1.0, np.max(np.abs(hist["eloc"][0])))
        print(f"  identical samples, E_loc rel err {err:.2e}", flush=True)
        assert err < 1e-9
    else:
        # complex64 contractions: E_loc of the configurations both chains hold
        same = np.all(cfg.cpu().numpy() == hist["samples"][0], axis=1)
        err = np.abs(e - hist["eloc"][0])[same] / np.maximum(1.0, np.abs(hist["eloc"][0][same]))
        print(f"  {same.sum()} of {n_chains} chains agree, E_loc rel err max {err.max():.2e}", flush=True)
        assert same.sum() >= n_chains // 2 and err.max() < 1e-3
