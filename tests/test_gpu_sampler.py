"""GPU parity of vmc_init_cache / vmc_mc_sweep against the oracle sampler.

The device and the oracle draw from the same Philox4x32-10 stream (counter
(site, sweep, chain, 0)), so in FP64 the sampled configurations must be
identical and amp, o, p and E_loc must agree to rounding.
Oracle: oracle.peps.mc_sampler (src/vmc/core/sampling.py:22-64 over
transition/estimate of src/vmc/peps/standard/kernels.py:131-245).
"""
import numpy as np
import pytest
import torch

from oracle import hamiltonians as HH
from oracle import peps as P

pytestmark = pytest.mark.gpu


def _tables(terms, shape, d=2):
    R, C = shape
    t = {"diag_sites": [], "diag_vals": [], "one_sites": [], "one_ops": [], "h_sites": [], "h_ops": [],
         "v_sites": [], "v_ops": []}
    for sites, diag in terms.diagonal:
        fl = [r * C + c for r, c in sites]
        t["diag_sites"] += [fl[0], fl[1] if len(fl) > 1 else -1]
        v = np.zeros(d * d)
        v[: len(diag)] = diag
        t["diag_vals"].append(v)
    for r in range(R):
        for c in range(C):
            for op in terms.one_site[r][c]:
                t["one_sites"].append(r * C + c)
                t["one_ops"].append(np.asarray(op, float).ravel())
            if c + 1 < C:
                for op in terms.horizontal[r][c]:
                    t["h_sites"].append(r * C + c)
                    t["h_ops"].append(np.asarray(op, float).ravel())
            if r + 1 < R:
                for op in terms.vertical[r][c]:
                    t["v_sites"].append(r * C + c)
                    t["v_ops"].append(np.asarray(op, float).ravel())
    return {k: (np.asarray(v, np.int32) if k.endswith("sites") else
                (np.concatenate(v) if v else np.zeros(0))) for k, v in t.items()}


def _run(shape, D, chi, terms, dtype, n_chains, n_steps, seed):
    from vmc_amd.engine import AmpGradPlan

    print(f"[sampler] shape={shape} D={D} chi={chi} dtype={dtype} chains={n_chains} steps={n_steps}", flush=True)
    T = P.random_peps(shape, D, seed=seed)
    flat = np.concatenate([t.ravel() for row in T for t in row])
    cfg0 = P.random_configs(shape[0] * shape[1], n_chains, seed=seed + 17)
    plan = AmpGradPlan(shape, D, chi, dtype=dtype)
    plan.set_terms(_tables(terms, shape))
    params = torch.tensor(flat, dtype=dtype, device="cuda")
    cfg = torch.tensor(cfg0, dtype=torch.int32, device="cuda")
    cache = plan.new_cache(n_chains)
    plan.init_cache(params, cfg, cache)
    out = []
    for t in range(n_steps):
        la, o, p, e = plan.mc_sweep(params, cfg, cache, seed, t)
        out.append((cfg.cpu().numpy().copy(), la.cpu().numpy(), o.double().cpu().numpy(), p.cpu().numpy(),
                    e.cpu().numpy()))
    torch.cuda.synchronize()
    strat = P.ZipUp(chi)
    cache_ref = P.init_cache(T, cfg0, shape, strat)
    _, _, hist = P.mc_sampler(T, cfg0, cache_ref, shape, strat, terms, seed, n_steps)
    return out, hist


def _check(shape, D, chi, terms, dtype, n_chains, n_steps, seed, tol):
    out, hist = _run(shape, D, chi, terms, dtype, n_chains, n_steps, seed)
    for t, (cfg, la, o, p, e) in enumerate(out):
        assert np.array_equal(cfg, hist["samples"][t]), f"samples differ at sweep {t}"
        amp = hist["amp"][t]
        err_la = np.max(np.abs(la[:, 0] - np.log(np.abs(amp))))
        o_ref = hist["o"][t]
        err_o = np.max(np.abs(o - o_ref)) / np.max(np.abs(o_ref))
        e_ref = hist["eloc"][t]
        err_e = np.max(np.abs(e - e_ref)) / max(1.0, np.max(np.abs(e_ref)))
        print(f"  sweep {t}: |dlog|={err_la:.2e} o={err_o:.2e} eloc={err_e:.2e}", flush=True)
        assert np.array_equal(p, hist["p"][t])
        assert err_la < tol and err_o < tol * 10 and err_e < tol * 10


def test_sampler_tfim_4x4_f64():
    _check((4, 4), 2, 4, HH.tfim((4, 4), 1.0, 3.04), torch.float64, 16, 3, 11, 1e-10)


def test_sampler_heisenberg_3x3_f64():
    _check((3, 3), 2, 4, HH.heisenberg((3, 3)), torch.float64, 16, 3, 5, 1e-10)


def test_sampler_heisenberg_4x4_cfg1_f64():
    _check((4, 4), 2, 4, HH.heisenberg((4, 4)), torch.float64, 8, 2, 0, 1e-10)


def test_sampler_heisenberg_4x4_f32():
    _check((4, 4), 2, 4, HH.heisenberg((4, 4)), torch.float32, 8, 2, 0, 2e-5)


def test_sampler_tfim_6x6_D3_f64():
    _check((6, 6), 3, 9, HH.tfim((6, 6), 1.0, 3.04), torch.float64, 4, 2, 2, 1e-9)
