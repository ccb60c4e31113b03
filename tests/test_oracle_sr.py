"""Pins the oracle's QGT / SR / integrator restatement (oracle.sr).

Recipes: tests/test_qgt.py (SliceOrdering blocks re-ordered == full Jacobian,
SliceOrdering solve == SiteOrdering solve), tests/test_qgt_solvers.py (solvers
vs a dense solve), tests/test_integrators.py (RK4 vs expm, Euler first order),
plus the sample/parameter-space duality of the regularised solve.
"""
import numpy as np
import pytest
import scipy.linalg as sla

from oracle import peps as P
from oracle import sr as S


def _sliced(seed=0, ns=64):
    shape, D = (2, 3), 2
    T = P.random_peps(shape, D, seed=seed)
    cfgs = P.random_configs(6, ns, seed=seed)
    amp, grad, p = P.value_and_grad(T, cfgs, shape, P.NoTruncation(), D)
    _, gfull, _ = P.value_and_grad(T, cfgs, shape, P.NoTruncation(), D, full_gradient=True)
    pps = P.params_per_site(shape, D)
    return grad / amp[:, None], p, gfull / amp[:, None], pps, (2,) * 6


def _sliced_complex(seed=0, ns=64):
    shape, D = (2, 3), 2
    T = P.random_peps(shape, D, seed=seed, dtype=np.complex128)
    cfgs = P.random_configs(6, ns, seed=seed)
    amp, grad, p = P.value_and_grad(T, cfgs, shape, P.NoTruncation(), D)
    _, gfull, _ = P.value_and_grad(T, cfgs, shape, P.NoTruncation(), D, full_gradient=True)
    pps = P.params_per_site(shape, D)
    return grad / amp[:, None], p, gfull / amp[:, None], pps, (2,) * 6


def test_complex_solve_equals_regularised_least_squares():
    """(O_c^H O_c / Ns + eps) x = O_c^H dv is the normal equation of
    min |[O_c/sqrt(Ns); sqrt(eps) I] x - [sqrt(Ns) dv; 0]|, solved here by
    numpy's SVD least squares: a dropped conjugate, a transposed operand or a
    wrong grad-factor phase in the restatement fails this."""
    o, p, ofull, pps, sd = _sliced_complex(6)
    assert np.iscomplexobj(ofull) and np.abs(ofull.imag).max() > 0.1 * np.abs(ofull).max()
    rng = np.random.default_rng(7)
    e = rng.standard_normal(len(o)) + 1j * rng.standard_normal(len(o))
    eps, ns = 1e-3, len(o)
    Oc = ofull - ofull.mean(axis=0)
    Sm = S.qgt_dense_parameter(ofull)
    assert np.allclose(Sm, Sm.conj().T, atol=1e-12 * np.abs(Sm).max())
    assert np.allclose(Sm, Oc.conj().T @ Oc / ns, atol=1e-12 * np.abs(Sm).max())
    A = np.vstack([Oc / np.sqrt(ns), np.sqrt(eps) * np.eye(Oc.shape[1])])
    for gf in (-1.0j, -1.0):
        dv = gf * (e - e.mean()) / ns
        b = np.concatenate([np.sqrt(ns) * dv, np.zeros(Oc.shape[1])])
        ref = np.linalg.lstsq(A, b, rcond=None)[0]
        for space in ("parameter", "sample"):
            x, _ = S.sr_apply(o, p, e, pps, sd, space=space, diag_shift=eps, grad_factor=gf, real_params=False)
            assert np.allclose(x, ref, rtol=1e-7, atol=1e-9 * np.abs(ref).max()), (gf, space)


def test_slice_blocks_reordered_equal_full_jacobian():
    o, p, ofull, pps, sd = _sliced()
    O = S.dense_blocks(o, p, sd, S.SliceOrdering())
    assert O.shape == ofull.shape
    perm_cols = S.reorder_slice_to_site(np.arange(O.shape[1]), pps, sd)
    assert np.array_equal(O[:, perm_cols], ofull)
    Osite = S.dense_blocks(o, p, sd, S.SiteOrdering(tuple(pps)))
    assert np.array_equal(Osite, ofull)


def test_qgt_hermitian_psd_and_matvec():
    o, p, ofull, pps, sd = _sliced(1)
    Sm = S.qgt_dense_parameter(ofull)
    Tm = S.qgt_dense_sample(ofull)
    assert np.allclose(Sm, Sm.conj().T, atol=1e-12 * np.abs(Sm).max())
    assert np.linalg.eigvalsh(Sm).min() > -1e-10 * np.abs(Sm).max()
    assert np.linalg.eigvalsh(Tm).min() > -1e-10 * np.abs(Tm).max()
    v = np.random.default_rng(0).standard_normal(Sm.shape[0])
    assert np.allclose(S.qgt_matvec_parameter(ofull, v), Sm @ v, rtol=1e-10, atol=1e-12)
    w = np.random.default_rng(1).standard_normal(Tm.shape[0])
    assert np.allclose(S.qgt_matvec_sample(ofull, w), Tm @ w, rtol=1e-10, atol=1e-12)
    # centred forms: S = Oc^H Oc / Ns, T = Oc Oc^H / Ns
    Oc = ofull - ofull.mean(axis=0)
    assert np.allclose(Sm, Oc.conj().T @ Oc / len(Oc), atol=1e-12 * np.abs(Sm).max())
    assert np.allclose(Tm, Oc @ Oc.conj().T / len(Oc), atol=1e-12 * np.abs(Tm).max())


def test_sample_and_parameter_space_solutions_agree():
    o, p, ofull, pps, sd = _sliced(2)
    e = np.random.default_rng(3).standard_normal(len(o))
    for gf in (1.0, -1.0j):
        xp, _ = S.sr_apply(o, p, e, pps, sd, space="parameter", diag_shift=1e-3, grad_factor=gf)
        xs, _ = S.sr_apply(o, p, e, pps, sd, space="sample", diag_shift=1e-3, grad_factor=gf)
        assert np.allclose(xp, xs, rtol=1e-7, atol=1e-9 * np.abs(xp).max())
    xso, _ = S.sr_apply(o, p, e, pps, sd, space="parameter", diag_shift=1e-3,
                        ordering=S.SiteOrdering(tuple(pps)))
    xsl, _ = S.sr_apply(o, p, e, pps, sd, space="parameter", diag_shift=1e-3)
    assert np.allclose(xso, xsl, rtol=1e-9, atol=1e-12)


def test_solvers_agree_with_dense_solve():
    rng = np.random.default_rng(4)
    A = rng.standard_normal((30, 30))
    A = A @ A.T + 30 * np.eye(30)
    b = rng.standard_normal(30)
    ref = np.linalg.solve(A, b)
    assert np.allclose(S.solve_cholesky(A, b), ref, rtol=1e-10)
    assert np.allclose(S.solve_svd(A, b), ref, rtol=1e-8)
    assert np.allclose(S.solve_cg(A, b, tol=1e-12), ref, rtol=1e-6)


def test_qr_solve_full_rank_matches_least_squares():
    rng = np.random.default_rng(5)
    O = rng.standard_normal((80, 12))
    dv = rng.standard_normal(80)
    x, m = S.qr_solve(O, dv)
    Oc = O - O.mean(axis=0)
    ref = np.linalg.lstsq(Oc, dv, rcond=None)[0]
    assert m["rank"] == 12
    assert np.allclose(x, ref, rtol=1e-9)


@pytest.mark.parametrize("complex_", [False, True])
def test_qr_solve_rank_deficient(complex_):
    """Ns < n: the centred Jacobian has rank Ns - 1, R22 vanishes, and the
    min-norm branch must return the minimum-norm least-squares solution
    pinv(O_c) dv; the basic branch (min_norm=False) must be a least-squares
    solution with at most rank nonzero entries."""
    rng = np.random.default_rng(11)
    ns, n = 24, 60
    O = rng.standard_normal((ns, n))
    dv = rng.standard_normal(ns)
    if complex_:
        O = O + 1j * rng.standard_normal((ns, n))
        dv = dv + 1j * rng.standard_normal(ns)
    Oc = O - O.mean(axis=0)
    x, m = S.qr_solve(O, dv)
    assert m["rank"] == ns - 1
    ref = np.linalg.pinv(Oc, rcond=1e-10) @ dv
    assert np.allclose(x, ref, rtol=1e-9, atol=1e-10 * np.abs(ref).max())
    xb, mb = S.qr_solve(O, dv, min_norm=False)
    assert np.count_nonzero(np.abs(xb) > 0) <= mb["rank"]
    g = Oc.conj().T @ (Oc @ xb - dv)
    assert np.linalg.norm(g) < 1e-9 * np.linalg.norm(Oc.conj().T @ dv)


def test_rk4_matches_expm_and_euler_first_order():
    rng = np.random.default_rng(0)
    M = rng.standard_normal((4, 4)) + 1j * rng.standard_normal((4, 4))
    M = 0.1 * (M + M.conj().T)
    y0 = rng.standard_normal(4) + 1j * rng.standard_normal(4)
    f = lambda y, t: M @ y
    y1, t1 = S.rk4_step(f, y0, 0.0, 0.1)
    ex = sla.expm(M * 0.1) @ y0
    assert np.linalg.norm(y1 - ex) / np.linalg.norm(ex) < 1e-7
    assert t1 == 0.1
    Mr = 0.05 * (lambda a: a + a.T)(rng.standard_normal((3, 3)))
    y0r = rng.standard_normal(3)
    ye, _ = S.euler_step(lambda y, t: Mr @ y, y0r, 0.0, 0.1)
    exr = sla.expm(Mr * 0.1) @ y0r
    assert np.linalg.norm(ye - exr) / np.linalg.norm(exr) < 1e-3
