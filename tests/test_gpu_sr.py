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


def _data(shape=(3, 3), D=2, ns=96, seed=0, complex_=False):
    T = P.random_peps(shape, D, seed=seed, dtype=np.complex128 if complex_ else np.float64)
    cfg = P.random_configs(shape[0] * shape[1], ns, seed=seed)
    amp, grad, p = P.value_and_grad(T, cfg, shape, P.NoTruncation(), D)
    o = grad / amp[:, None]
    rng = np.random.default_rng(seed)
    e = rng.standard_normal(ns)
    if complex_:
        e = e + 1j * rng.standard_normal(ns)
    pps = P.params_per_site(shape, D)
    return o, p, e, pps


def _model(shape, D, dtype=torch.float64):
    from vmc_amd.peps import PEPS, ZipUp

    return PEPS(shape=shape, bond_dim=D, contraction_strategy=ZipUp(D * D), dtype=dtype)


def _tfim_op(shape, J, h):
    from vmc_amd.operators import DiagonalTerm, LocalHamiltonian, OneSiteTerm

    terms = []
    for r in range(shape[0]):
        for c in range(shape[1]):
            terms.append(OneSiteTerm(r, c, -h * np.array([[0.0, 1.0], [1.0, 0.0]])))
            if c + 1 < shape[1]:
                terms.append(DiagonalTerm(((r, c), (r, c + 1)), -J * np.array([1.0, -1, -1, 1])))
            if r + 1 < shape[0]:
                terms.append(DiagonalTerm(((r, c), (r + 1, c)), -J * np.array([1.0, -1, -1, 1])))
    return LocalHamiltonian(shape=shape, terms=tuple(terms))


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


def _synthetic(ns, shape=(3, 3), D=3, seed=11, complex_=False):
    """Random log-derivatives and slice indices (the SR layer takes o and p as
    given), sized to span several 128-wide GEMM tiles with ragged tails and
    several 64-row triangular-solve blocks."""
    pps = P.params_per_site(shape, D)
    rng = np.random.default_rng(seed)
    o = rng.standard_normal((ns, sum(pps)))
    e = rng.standard_normal(ns)
    if complex_:
        o = o + 1j * rng.standard_normal(o.shape)
        e = e + 1j * rng.standard_normal(ns)
    p = rng.integers(0, 2, size=o.shape).astype(np.int8)
    return o, p, e, pps


@pytest.mark.parametrize("space,ns", [("sample", 300), ("parameter", 150), ("sample", 64)])
def test_sr_solve_multi_tile(space, ns):
    from vmc_amd.preconditioners import SRPreconditioner
    from vmc_amd.qgt import ParameterSpace, SampleSpace

    print(f"[sr] multi-tile space={space} ns={ns}", flush=True)
    o, p, e, pps = _synthetic(ns)
    ref, ref_m = S.sr_apply(o, p, e, pps, (2,) * 9, space=space, diag_shift=1e-3, grad_factor=-1.0)
    pre = SRPreconditioner(ParameterSpace() if space == "parameter" else SampleSpace(), diag_shift=1e-3)
    x, m = pre.apply(_model((3, 3), 3), None, None, torch.tensor(o, device="cuda"), torch.tensor(p, device="cuda"),
                     torch.tensor(e, device="cuda"), grad_factor=-1.0)
    x = x.cpu().numpy()
    err = np.max(np.abs(x - ref)) / np.max(np.abs(ref))
    print(f"  rel err {err:.2e} residual {m['residual_error']:.2e} (oracle {ref_m['residual_error']:.2e})", flush=True)
    assert err < 1e-9
    assert abs(m["residual_error"] - ref_m["residual_error"]) < 1e-6 * max(1.0, ref_m["residual_error"]) + 1e-10


@pytest.mark.parametrize("space,ns", [("sample", 600), ("parameter", 300)])
def test_sr_solve_complex_multi_panel(space, ns):
    """Complex128 solves large enough for the matrix-core products (four real
    MFMA products per complex one) and a Cholesky of more than one 256-column
    panel in sample space (its lower-triangle trailing update), real-time
    grad factor -i."""
    from vmc_amd.preconditioners import SRPreconditioner
    from vmc_amd.qgt import ParameterSpace, SampleSpace

    print(f"[sr] complex multi-panel space={space} ns={ns}", flush=True)
    o, p, e, pps = _synthetic(ns, complex_=True, seed=13)
    ref, ref_m = S.sr_apply(o, p, e, pps, (2,) * 9, space=space, diag_shift=1e-3, grad_factor=-1.0j,
                            real_params=False)
    pre = SRPreconditioner(ParameterSpace() if space == "parameter" else SampleSpace(), diag_shift=1e-3)
    x, m = pre.apply(_model((3, 3), 3, torch.complex128), None, None, torch.tensor(o, device="cuda"),
                     torch.tensor(p, device="cuda"), torch.tensor(e, device="cuda"), grad_factor=-1.0j)
    x = x.cpu().numpy()
    err = np.max(np.abs(x - ref)) / np.max(np.abs(ref))
    print(f"  rel err {err:.2e} residual {m['residual_error']:.2e} (oracle {ref_m['residual_error']:.2e})", flush=True)
    assert err < 1e-9
    assert abs(m["residual_error"] - ref_m["residual_error"]) < 1e-6 * max(1.0, ref_m["residual_error"]) + 1e-10


@pytest.mark.parametrize("complex_", [False, True])
def test_qgt_dense_multi_tile(complex_):
    """Both dense QGTs over several tiles: the real products fill the upper
    triangle by mirroring the lower one."""
    from vmc_amd.qgt import QGT, ParameterSpace, SampleSpace, SlicedJacobian

    print(f"[sr] qgt dense multi-tile complex={complex_}", flush=True)
    o, p, _, pps = _synthetic(200, complex_=complex_)
    jac = SlicedJacobian(torch.tensor(o, device="cuda"), torch.tensor(p, device="cuda"), (2,) * 9,
                         params_per_site=tuple(pps))
    O = S.dense_blocks(o, p, (2,) * 9, S.SiteOrdering(tuple(pps)))
    Sd = QGT(jac, ParameterSpace()).to_dense().cpu().numpy()
    Td = QGT(jac, SampleSpace()).to_dense().cpu().numpy()
    assert np.abs(Sd - S.qgt_dense_parameter(O)).max() < 1e-12 * np.abs(Sd).max()
    assert np.abs(Td - S.qgt_dense_sample(O)).max() < 1e-12 * np.abs(Td).max()


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


@pytest.mark.parametrize("space", ["parameter", "sample"])
@pytest.mark.parametrize("odt", [torch.complex128, torch.complex64])
def test_sr_solve_complex_real_time(space, odt):
    """Complex parameters, complex E_loc, grad factor -i (RealTimeUnit)."""
    from vmc_amd.preconditioners import SRPreconditioner
    from vmc_amd.qgt import ParameterSpace, SampleSpace

    print(f"[sr] complex space={space} o={odt}", flush=True)
    shape, D = (3, 3), 2
    o, p, e, pps = _data(shape, D, complex_=True, seed=4)
    o_dev = torch.tensor(o, device="cuda").to(odt)
    o_in = o_dev.cpu().numpy().astype(np.complex128)
    ref, ref_m = S.sr_apply(o_in, p, e, pps, (2,) * 9, space=space, diag_shift=1e-3, grad_factor=-1.0j,
                            real_params=False)
    pre = SRPreconditioner(ParameterSpace() if space == "parameter" else SampleSpace(), diag_shift=1e-3)
    x, m = pre.apply(_model(shape, D, torch.complex128), None, None, o_dev, torch.tensor(p, device="cuda"),
                     torch.tensor(e, device="cuda"), grad_factor=-1.0j)
    assert x.dtype == torch.complex128
    x = x.cpu().numpy()
    err = np.max(np.abs(x - ref)) / np.max(np.abs(ref))
    print(f"  rel err {err:.2e} residual {m['residual_error']:.2e} (oracle {ref_m['residual_error']:.2e})", flush=True)
    # same conditioning bound as the real case (shifted solve below the rank)
    assert err < (1e-7 if space == "parameter" else 1e-9)
    assert abs(m["residual_error"] - ref_m["residual_error"]) < 1e-6 * max(1.0, ref_m["residual_error"]) + 1e-10


def test_qgt_dense_complex_matches_oracle():
    from vmc_amd.qgt import QGT, ParameterSpace, SampleSpace, SlicedJacobian

    print("[sr] qgt dense complex", flush=True)
    o, p, e, pps = _data(complex_=True, seed=5)
    jac = SlicedJacobian(torch.tensor(o, device="cuda"), torch.tensor(p, device="cuda"), (2,) * 9,
                         params_per_site=tuple(pps))
    O = S.dense_blocks(o, p, (2,) * 9, S.SiteOrdering(tuple(pps)))
    Sd = QGT(jac, ParameterSpace()).to_dense().cpu().numpy()
    Td = QGT(jac, SampleSpace()).to_dense().cpu().numpy()
    assert np.allclose(Sd, S.qgt_dense_parameter(O), atol=1e-12 * np.abs(Sd).max())
    assert np.allclose(Td, S.qgt_dense_sample(O), atol=1e-12 * np.abs(Td).max())


def test_real_time_tdvp_derivative_matches_oracle():
    """One TDVP time derivative (tdvp.py:113-163) of a complex PEPS under
    RealTimeUnit against the oracle sampler + SR on the same Philox stream."""
    from vmc_amd.drivers import RealTimeUnit, TDVPDriver
    from vmc_amd.peps import PEPS, ZipUp
    from vmc_amd.preconditioners import SRPreconditioner
    from vmc_amd.qgt import SampleSpace
    from oracle.hamiltonians import tfim

    print("[sr] real-time tdvp derivative 3x3 complex", flush=True)
    shape, D, chi, ns, nc, seed = (3, 3), 2, 4, 60, 16, 5
    T = P.random_peps(shape, D, seed=11, dtype=np.complex128)
    model = PEPS(shape=shape, bond_dim=D, contraction_strategy=ZipUp(chi), tensors=T, dtype=torch.complex128)
    drv = TDVPDriver(model, _tfim_op(shape, 1.0, 3.04), preconditioner=SRPreconditioner(SampleSpace(), diag_shift=1e-3),
                     dt=0.01, time_unit=RealTimeUnit(), n_samples=ns, n_chains=nc, sampler_seed=seed)
    cfg0 = drv._configs.cpu().numpy().astype(np.int64)
    upd, _, (e, metrics) = drv._time_derivative(drv._params.clone(), 0.0, (drv._key, drv._configs))
    _, _, chain_length, total = P.sample_counts(ns, nc)
    cache = P.init_cache(T, cfg0, shape, P.ZipUp(chi))
    _, _, hist = P.mc_sampler(T, cfg0, cache, shape, P.ZipUp(chi), tfim(shape, 1.0, 3.04), seed, chain_length)
    o = P.trim_samples(hist["o"], total, ns)
    p = P.trim_samples(hist["p"], total, ns)
    er = P.trim_samples(hist["eloc"], total, ns)
    assert np.allclose(e.cpu().numpy(), er, rtol=1e-9, atol=1e-9)
    ref, _ = S.sr_apply(o, p, er, P.params_per_site(shape, D), (2,) * 9, space="sample", diag_shift=1e-3,
                        grad_factor=-1.0j, real_params=False)
    got = upd.cpu().numpy()
    err = np.max(np.abs(got - ref)) / np.max(np.abs(ref))
    print(f"  update rel err {err:.2e}", flush=True)
    assert err < 1e-8
    # two RK4 steps advance time and keep the state finite
    drv.run(0.02)
    assert drv.step_count == 2 and abs(drv.t - 0.02) < 1e-12
    assert torch.isfinite(torch.view_as_real(drv._params)).all()


@pytest.mark.parametrize("case", ["full_rank", "rank_deficient", "basic", "complex_real_time"])
def test_qr_solve_matches_oracle(case):
    """QRSolve (preconditioners.py:166-214) on the full Jacobian vs oracle.sr.qr_solve."""
    from vmc_amd.preconditioners import QRSolve, SRPreconditioner
    from vmc_amd.qgt import ParameterSpace

    print(f"[sr] QRSolve {case}", flush=True)
    complex_ = case == "complex_real_time"
    shape, D = ((2, 2), 2) if case == "full_rank" else ((3, 3), 2)
    ns = 96 if case == "full_rank" else 40
    T = P.random_peps(shape, D, seed=3, dtype=np.complex128 if complex_ else np.float64)
    cfg = P.random_configs(shape[0] * shape[1], ns, seed=3)
    amp, grad, _ = P.value_and_grad(T, cfg, shape, P.NoTruncation(), D, full_gradient=True)
    O = grad / amp[:, None]
    rng = np.random.default_rng(3)
    e = rng.standard_normal(ns) + (1j * rng.standard_normal(ns) if complex_ else 0.0)
    gf = -1.0j if complex_ else -1.0
    dv = gf * (e - e.mean()) / ns
    ref, ref_m = S.qr_solve(O, dv, min_norm=case != "basic")
    if not complex_:
        ref = 2.0 * ref.real
    pre = SRPreconditioner(ParameterSpace(), QRSolve(min_norm=case != "basic"))
    model = _model(shape, D, torch.complex128 if complex_ else torch.float64)
    x, m = pre.apply(model, None, None, torch.tensor(O, device="cuda"), None, torch.tensor(e, device="cuda"),
                     grad_factor=gf)
    x = x.cpu().numpy()
    err = np.max(np.abs(x - ref)) / np.max(np.abs(ref))
    print(f"  rank {m['rank']} (oracle {ref_m['rank']}) rel err {err:.2e} residual {m['residual_error']:.2e} "
          f"(oracle {ref_m['residual_error']:.2e})", flush=True)
    assert m["rank"] == ref_m["rank"]
    assert err < 1e-8


def test_sr_solve_large_vs_torch():
    """Sample-space solve at 4096 samples x 56448 expanded columns (12x12 D=4
    site-major layout: 32 GEMM tile rows, 16 Cholesky panels of 256, 64
    triangular-solve blocks) against torch FP64 (rocBLAS product, rocSOLVER
    Cholesky) on the same random log-derivatives."""
    from vmc_amd.preconditioners import SRPreconditioner
    from vmc_amd.qgt import SampleSpace

    shape, D, ns, shift = (12, 12), 4, 4096, 1e-4
    print(f"[sr] large sample-space solve ns={ns} vs torch", flush=True)
    pps = P.params_per_site(shape, D)
    g = torch.Generator(device="cuda").manual_seed(0)
    o = torch.randn((ns, sum(pps)), dtype=torch.float64, device="cuda", generator=g)
    p = (torch.rand((ns, sum(pps)), device="cuda", generator=g) < 0.5).to(torch.int8)
    e = torch.randn(ns, dtype=torch.float64, device="cuda", generator=g)
    pre = SRPreconditioner(SampleSpace(), diag_shift=shift)
    x, m = pre.apply(_model(shape, D), None, None, o, p, e, grad_factor=-1.0)
    cols, off = [], 0
    for n in pps:
        for k in range(2):
            cols.append(torch.where(p[:, off:off + n] == k, o[:, off:off + n], torch.zeros((), dtype=o.dtype,
                                                                                           device=o.device)))
        off += n
    O = torch.cat(cols, dim=1)
    del cols
    O -= O.mean(dim=0, keepdim=True)
    dv = -(e - e.mean()) / ns
    T = (O @ O.T) / ns
    T.diagonal().add_(shift)
    y = torch.cholesky_solve(dv[:, None], torch.linalg.cholesky(T))[:, 0]
    xr = 2.0 * (O.T @ y)
    res = ((O @ (O.T @ y) / ns - dv).norm() ** 2 / dv.norm() ** 2).item()
    err = ((x - xr).abs().max() / xr.abs().max()).item()
    print(f"  rel err {err:.2e} residual {m['residual_error']:.3e} (torch {res:.3e})", flush=True)
    assert err < 1e-11
    assert abs(m["residual_error"] - res) < 1e-6 * res + 1e-14
