"""Autotuner (vmc_plan_autotune / vmc_plan_tuning / vmc_plan_set_tuning).

* A tuning pass records one choice per distinct truncation and theta shape of
  the plan, each a kernel that can run that shape, with the timings it was
  picked on; its outputs and every later call agree with the untuned plan.
* Every forced choice agrees with the default kernels: theta tiles bit for bit
  (every tile accumulates in the same K order), the 16-lane multisection and
  the generic truncation kernel to rounding (they compute the same subspace).
* A generic choice grows the workspace; the sampler, density-matrix, complex,
  variational and gauge-invariant plans run with tuned choices.
* Tables round-trip between identical plans; choices that cannot run a shape
  are refused and leave the table unchanged.
"""
import numpy as np
import pytest
import torch

from oracle import peps as P

pytestmark = pytest.mark.gpu


def _flat(T):
    return np.concatenate([t.ravel() for row in T for t in row])


def _setup(shape=(6, 6), D=4, chi=16, dtype=torch.float64, strategy=None, B=128, seed=31):
    from vmc_amd import _lib
    from vmc_amd.engine import AmpGradPlan

    cplx = dtype.is_complex
    T = P.random_peps(shape, D, seed=seed, dtype=np.complex128 if cplx else np.float64)
    cfg = torch.tensor(P.random_configs(shape[0] * shape[1], B, seed=seed + 1))
    plan = AmpGradPlan(shape, D, chi, dtype=dtype, strategy=_lib.VMC_ZIPUP if strategy is None else strategy)
    params = torch.tensor(_flat(T), dtype=dtype, device="cuda")
    return plan, params, cfg


def _run(plan, params, cfg):
    la, o, _ = plan.amp_grad(params, cfg)
    torch.cuda.synchronize()
    return la.cpu().numpy(), o.cpu().numpy()


def _close(a, b, tol_log=1e-12, tol_o=1e-9):
    (la0, o0), (la1, o1) = a, b
    assert np.max(np.abs(la1[:, 0] - la0[:, 0]) / np.maximum(1.0, np.abs(la0[:, 0]))) < tol_log
    assert np.max(np.abs(np.angle(np.exp(1j * (la1[:, 1] - la0[:, 1]))))) < 1e3 * tol_log
    assert np.max(np.abs(o1 - o0)) <= tol_o * np.max(np.abs(o0))


def _force(entries, kind, choice):
    return [dict(e, choice=choice) if e["kind"] == kind else e for e in entries]


def test_autotune_records_legal_choices_and_keeps_results():
    from vmc_amd import _lib

    plan, params, cfg = _setup()
    print("[autotune] 6x6 D4 chi16 f64 B=128", flush=True)
    ref = _run(plan, params, cfg)
    ws0 = plan.workspace(cfg.shape[0]).numel()
    entries = plan.autotune(params, cfg)
    kinds = [e["kind"] for e in entries]
    print(f"  {kinds.count('trunc')} truncation and {kinds.count('theta')} theta shapes:", flush=True)
    for e in entries:
        print(f"    {e}", flush=True)
    # cuts that keep every column (k = n < m) run the QR kernel and have no entry
    assert kinds.count("trunc") >= 1 and kinds.count("theta") >= 4
    assert len({(e["kind"], e["key"]) for e in entries}) == len(entries)
    for e in entries:
        assert e["ns"] > 0 and e["default_ns"] > 0
        assert e["ns"] <= e["default_ns"]  # a non-default choice beat the default
        if e["kind"] == "trunc":
            m, n, k, dm = e["key"]
            assert dm == 0 and k < m and m <= 128 and n <= 128 and k < n
            assert e["choice"] in (_lib.VMC_TRUNC_LPE8, _lib.VMC_TRUNC_LPE16, _lib.VMC_TRUNC_GENERIC)
        else:
            assert 0 <= e["choice"] < 16
    _close(ref, _run(plan, params, cfg))
    if any(e["kind"] == "trunc" and e["choice"] == _lib.VMC_TRUNC_GENERIC for e in entries):
        assert plan.workspace(cfg.shape[0]).numel() > ws0


def _factored_rows(key):
    """Rows the truncation kernel factors: n on the theta^T route (k < n < m, real)."""
    m, n, k, dm = key
    return n if (not dm and k < n < m and m <= 128 and k <= 32) else m


def test_forced_choices_agree_with_the_defaults():
    from vmc_amd import _lib

    plan, params, cfg = _setup(chi=32)
    print("[autotune] forced choices 6x6 D4 chi32 f64", flush=True)
    ref = _run(plan, params, cfg)
    entries = plan.autotune(params, cfg, reps=1)
    base = _force(_force(entries, "trunc", _lib.VMC_TRUNC_LPE8), "theta", 0)
    plan.set_tuning(base)
    assert _run(plan, params, cfg)[0].tobytes() == ref[0].tobytes()
    for t1 in range(4):
        for t2 in range(4):
            plan.set_tuning(_force(base, "theta", t1 + 4 * t2))
            la, o = _run(plan, params, cfg)
            assert la.tobytes() == ref[0].tobytes() and o.tobytes() == ref[1].tobytes(), (t1, t2)
    ws0 = plan.workspace(cfg.shape[0]).numel()
    # the 16-lane variant exists for cuts factoring more than 64 rows only
    lpe16 = [dict(e, choice=_lib.VMC_TRUNC_LPE16) if e["kind"] == "trunc" and _factored_rows(e["key"]) > 64 else e
             for e in base]
    assert lpe16 != base
    for v, table in ((_lib.VMC_TRUNC_LPE16, lpe16), (_lib.VMC_TRUNC_GENERIC, _force(base, "trunc", _lib.VMC_TRUNC_GENERIC))):
        plan.set_tuning(table)
        got = _run(plan, params, cfg)
        err = np.max(np.abs(got[0][:, 0] - ref[0][:, 0]) / np.abs(ref[0][:, 0]))
        print(f"  truncation variant {v}: max rel |dlog psi| {err:.2e}", flush=True)
        _close(ref, got)
    assert plan.workspace(cfg.shape[0]).numel() > ws0  # generic scratch for every cut
    plan.set_tuning([])
    assert plan.tuning() == []
    assert plan.workspace(cfg.shape[0]).numel() == ws0


def test_tables_round_trip_and_invalid_entries_are_refused():
    from vmc_amd import _lib

    plan, params, cfg = _setup((4, 4), 3, 6, B=64)
    print("[autotune] round trip 4x4 D3 chi6", flush=True)
    entries = plan.autotune(params, cfg, reps=1)
    other, _, _ = _setup((4, 4), 3, 6, B=64)
    other.set_tuning(entries)
    assert other.tuning() == entries
    trunc = next(e for e in entries if e["kind"] == "trunc")
    m, n, k, dm = trunc["key"]
    bad = [
        [dict(trunc, key=(m, n, m, dm))],                 # nothing discarded: no kernel choice
        [dict(trunc, choice=7)],                          # unknown variant
        [dict(trunc, key=(600, n, k, dm), choice=_lib.VMC_TRUNC_LPE8)],  # beyond the register kernel
        [dict(trunc, key=(64, 16, 16, 0), choice=_lib.VMC_TRUNC_LPE8)],  # every column kept: the QR kernel
        [dict(trunc, key=(64, 64, 16, 0), choice=_lib.VMC_TRUNC_LPE16)],  # 16 lanes: only beyond 64 rows
        [dict(next(e for e in entries if e["kind"] == "theta"), choice=16)],
    ]
    for b in bad:
        with pytest.raises(RuntimeError):
            other.set_tuning(b)
        assert other.tuning() == entries
    cplan, _, _ = _setup((4, 4), 3, 6, dtype=torch.complex128, B=64)
    with pytest.raises(RuntimeError, match="does not run"):
        cplan.set_tuning([dict(trunc, choice=_lib.VMC_TRUNC_LPE16)])  # real-only variant


@pytest.mark.parametrize("dtype", [torch.complex128, torch.float32])
def test_autotune_other_dtypes(dtype):
    from vmc_amd import _lib

    plan, params, cfg = _setup((6, 6), 3, 9, dtype=dtype, B=96)
    print(f"[autotune] 6x6 D3 chi9 {dtype}", flush=True)
    ref = _run(plan, params, cfg)
    entries = plan.autotune(params, cfg)
    for e in entries:
        if e["kind"] == "trunc" and dtype.is_complex:
            assert e["choice"] != _lib.VMC_TRUNC_LPE16
    if dtype == torch.float32:  # FP32 contractions: rounding of the theta inputs differs by kernel only in FP64 cuts
        _close(ref, _run(plan, params, cfg), tol_log=1e-5, tol_o=1e-3)
    else:
        _close(ref, _run(plan, params, cfg))


@pytest.mark.parametrize("strategy", ["density", "variational"])
def test_autotune_other_strategies(strategy):
    from vmc_amd import _lib

    code = {"density": _lib.VMC_DENSITY_MATRIX, "variational": _lib.VMC_VARIATIONAL}[strategy]
    # density-matrix cuts keep min(chi, n) eigenvectors, which can span null
    # space of theta^H theta whose basis is arbitrary and moves the amplitude;
    # chi <= D keeps the result basis independent
    chi = 3 if strategy == "density" else 9
    plan, params, cfg = _setup((6, 6), 3, chi, strategy=code, B=96)
    print(f"[autotune] 6x6 D3 chi{chi} {strategy}", flush=True)
    ref = _run(plan, params, cfg)
    entries = plan.autotune(params, cfg)
    trunc = [e for e in entries if e["kind"] == "trunc"]
    if strategy == "density":
        assert trunc and all(e["key"][3] == 1 for e in trunc)
    else:
        assert not trunc  # variational cuts run the QR / ALS kernels
    _close(ref, _run(plan, params, cfg))
    plan.set_tuning(_force(entries, "trunc", _lib.VMC_TRUNC_GENERIC))
    _close(ref, _run(plan, params, cfg))


def test_sampler_after_tuning():
    """init_cache + one estimate pass with every cut forced to the generic
    kernel (larger workspace) gives the untuned log psi and E_loc."""
    from vmc_amd import _lib
    from vmc_amd.operators import DiagonalTerm, LocalHamiltonian, OneSiteTerm
    from vmc_amd.operators.local_terms import to_tables

    shape = (6, 6)
    plan, params, cfg = _setup(shape, 3, 9, B=64)
    print("[autotune] sampler pass after tuning 6x6 D3 chi9", flush=True)
    zz = -np.array([1.0, -1.0, -1.0, 1.0])
    ox = -3.0 * np.array([[0.0, 1.0], [1.0, 0.0]])
    terms = [OneSiteTerm(r, c, ox) for r in range(6) for c in range(6)]
    terms += [DiagonalTerm(((r, c), (r, c + 1)), zz) for r in range(6) for c in range(5)]
    terms += [DiagonalTerm(((r, c), (r + 1, c)), zz) for r in range(5) for c in range(6)]
    plan.set_terms(to_tables(LocalHamiltonian(shape=shape, terms=tuple(terms)), 2))
    cfg = cfg.to(device="cuda", dtype=torch.int32)

    def est():
        la, e = plan.local_estimate(params, cfg)
        torch.cuda.synchronize()
        return la.cpu().numpy(), e.cpu().numpy()

    la0, e0 = est()
    entries = plan.autotune(params, cfg, reps=1)
    plan.set_tuning(_force(entries, "trunc", _lib.VMC_TRUNC_GENERIC))
    la1, e1 = est()
    assert np.max(np.abs(la1[:, 0] - la0[:, 0]) / np.abs(la0[:, 0])) < 1e-12
    assert np.max(np.abs(e1 - e0)) < 1e-8 * max(1.0, np.max(np.abs(e0)))


def test_autotune_gauge_invariant_plan():
    from oracle import gi as G
    from vmc_amd.engine import GIPlan

    shape, N, dmax, chi = (6, 6), 2, 2, 8
    print("[autotune] GI 6x6 N2 dmax2 chi8", flush=True)
    cfg = G.GIConfig(shape=shape, N=N, phys_dim=1, Qx=0, degeneracy_per_charge=(dmax,) * N, charge_of_site=(0,))
    T = G.random_gi_tensors(cfg, seed=5)
    s = torch.tensor(G.random_configurations(cfg, 64, seed=6))
    plan = GIPlan(shape, N, 1, dmax, chi, cfg.charge_of_site, dtype=torch.float64)
    params = torch.tensor(_flat(T), dtype=torch.float64, device="cuda")
    ref = _run(plan, params, s)
    entries = plan.autotune(params, s)
    assert any(e["kind"] == "trunc" for e in entries)
    _close(ref, _run(plan, params, s))
