"""GPU parity of the density-matrix strategy (strategy.py:153-176) and of
zip-up at truncating chi against closed forms and the FP64 oracle.

* The designed 2 x 3 PEPS of oracle.designed has monomial thetas in its top
  row, so the amplitudes of ZipUp(chi) and DensityMatrix(chi) are known in
  closed form through two truncating cuts, and the two rules keep different
  paths. Evaluated through vmc_amp_grad (amplitudes only).
* Random PEPS with chi <= D (no cut keeps null-space vectors, so the
  reference's density-matrix result is basis independent): log psi and o
  against oracle.peps.DensityMatrix.
* The truncated env-gradient row identity at the metric shape (10 x 10, D=4,
  chi=32, 4096 samples): sum over a row's sites of A[s] * o[s] is the same for
  every site of the row and equals 1 on the last row.
"""
import numpy as np
import pytest
import torch

from oracle import designed as DS
from oracle import peps as P

pytestmark = pytest.mark.gpu


def _flat(tensors):
    return np.concatenate([t.ravel() for row in tensors for t in row])


def _code(rule):
    from vmc_amd import _lib

    return {"zipup": _lib.VMC_ZIPUP, "density": _lib.VMC_DENSITY_MATRIX}[rule]


@pytest.mark.parametrize("dtype,tol", [(torch.float64, 1e-12), (torch.float32, 2e-6), (torch.complex128, 1e-12)])
@pytest.mark.parametrize("rule", ["zipup", "density"])
def test_designed_closed_form(rule, dtype, tol):
    from vmc_amd.engine import AmpGradPlan

    cfgs = DS.all_configs()
    for chi in (2, 3):
        plan = AmpGradPlan(DS.SHAPE, DS.D, chi, dtype=dtype, strategy=_code(rule))
        for seed in range(8):
            print(f"[dm] designed {rule} chi={chi} seed={seed} {dtype}", flush=True)
            T, w = DS.monomial_peps(seed)
            ref = DS.amplitudes(w, chi, rule, cfgs)
            params = torch.tensor(_flat(T), dtype=dtype, device="cuda")
            la, _, _ = plan.amp_grad(params, torch.tensor(cfgs), grad=False)
            la = la.cpu().numpy()
            sign = np.where(np.abs(la[:, 1]) > 1.0, -1.0, 1.0)  # phase pi or 0 (real weights)
            amp = sign * np.exp(la[:, 0])
            assert np.max(np.abs(amp - ref)) <= tol * np.max(np.abs(ref)), (seed, chi, amp, ref)


def _parity(shape, D, chi, dtype, B, seed, tol_log, tol_o):
    from vmc_amd.engine import AmpGradPlan

    cplx = dtype.is_complex
    T = P.random_peps(shape, D, seed=seed, dtype=np.complex128 if cplx else np.float64)
    cfg = P.random_configs(shape[0] * shape[1], B, seed=seed)
    print(f"[dm] parity {shape} D={D} chi={chi} {dtype} B={B}", flush=True)
    plan = AmpGradPlan(shape, D, chi, dtype=dtype, strategy=_code("density"))
    params = torch.tensor(_flat(T), dtype=dtype, device="cuda")
    la, o, p = plan.amp_grad(params, torch.tensor(cfg), want_p=True)
    la, o, p = la.cpu().numpy(), o.cpu().numpy(), p.cpu().numpy()
    amp, grad, p_ref = P.value_and_grad(T, cfg, shape, P.DensityMatrix(chi), D)
    ref_log = np.log(np.abs(amp))
    err_log = np.max(np.abs(la[:, 0] - ref_log) / np.maximum(1.0, np.abs(ref_log)))
    if cplx:
        dph = np.angle(np.exp(1j * (la[:, 1] - np.angle(amp))))
        assert np.max(np.abs(dph)) < 1e-9
    else:
        assert np.array_equal(np.where(la[:, 1] != 0, -1.0, 1.0), np.sign(amp))
    o_ref = grad / amp[:, None]
    err_o = np.max(np.abs(o - o_ref)) / np.max(np.abs(o_ref))
    print(f"  rel err log psi {err_log:.3e}  rel err o {err_o:.3e}", flush=True)
    assert err_log < tol_log and err_o < tol_o
    assert np.array_equal(p, p_ref)


def test_density_4x4_D3_chi3_f64():
    _parity((4, 4), 3, 3, torch.float64, 32, 21, 1e-12, 1e-9)


def test_density_10x10_D4_chi4_f64():
    _parity((10, 10), 4, 4, torch.float64, 16, 22, 1e-10, 1e-7)


def test_density_4x4_D3_chi3_c128():
    _parity((4, 4), 3, 3, torch.complex128, 16, 23, 1e-12, 1e-9)


def test_density_rejects_nothing_at_chi_above_D():
    """chi > D is accepted (the reference accepts it); its result depends on the
    null-space basis, so only finiteness is asserted here."""
    from vmc_amd.engine import AmpGradPlan

    shape, D, chi = (4, 4), 3, 4
    T = P.random_peps(shape, D, seed=24)
    cfg = P.random_configs(16, 16, seed=24)
    plan = AmpGradPlan(shape, D, chi, dtype=torch.float64, strategy=_code("density"))
    la, o, _ = plan.amp_grad(torch.tensor(_flat(T), device="cuda"), torch.tensor(cfg))
    assert torch.isfinite(la).all() and torch.isfinite(o).all()


@pytest.mark.parametrize("dtype,tol", [(torch.float64, 1e-9)])
def test_row_identity_metric_shape(dtype, tol):
    """sum_{site in row r} A[s] . o[s] = psi_r / psi, constant along row r; 1 on the last row."""
    from vmc_amd.engine import AmpGradPlan

    shape, D, chi, B = (10, 10), 4, 32, 4096
    print(f"[dm] row identity {shape} D={D} chi={chi} B={B} {dtype}", flush=True)
    T = P.random_peps(shape, D, seed=0)
    cfg = torch.tensor(P.random_configs(100, B, seed=1), device="cuda")
    plan = AmpGradPlan(shape, D, chi, dtype=dtype)
    params = torch.tensor(_flat(T), dtype=dtype, device="cuda")
    _, o, _ = plan.amp_grad(params, cfg)
    sums = torch.empty((B, 100), dtype=torch.float64, device="cuda")
    off_p, off_o = 0, 0
    for r in range(10):
        for c in range(10):
            t = T[r][c]
            n = t[0].size
            site = params[off_p:off_p + t.size].view(t.shape[0], n)
            a = site[cfg[:, r * 10 + c].long()]  # (B, n)
            sums[:, r * 10 + c] = (a.double() * o[:, off_o:off_o + n].double()).sum(1)
            off_p += t.size
            off_o += n
    sums = sums.view(B, 10, 10)
    dev = (sums - sums[:, :, :1]).abs().amax(dim=2) / sums[:, :, 0].abs()
    print(f"  max row spread {dev.max().item():.3e}, last row |s - 1| "
          f"{(sums[:, -1, :] - 1).abs().max().item():.3e}", flush=True)
    assert dev.max().item() < tol
    assert (sums[:, -1, :] - 1).abs().max().item() < tol
