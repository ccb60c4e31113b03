"""The amp+grad step captured in a HIP graph (AmpGradGraph) replays the eager
result bit for bit, several times with eager work and synchronisations in
between, and follows new inputs copied into its buffers."""
import numpy as np
import pytest
import torch

from oracle import peps as P

pytestmark = pytest.mark.gpu


def _flat(tensors):
    return np.concatenate([t.ravel() for row in tensors for t in row])


@pytest.mark.parametrize("shape,D,chi,dtype,B", [
    ((4, 4), 2, 4, torch.float64, 64),
    ((10, 10), 4, 32, torch.float32, 128),
    ((6, 6), 3, 9, torch.complex128, 32),
])
def test_captured_amp_grad_replays_eager(shape, D, chi, dtype, B):
    from vmc_amd.engine import AmpGradGraph, AmpGradPlan

    print(f"[graph] {shape} D={D} chi={chi} {dtype} B={B}", flush=True)
    cplx = dtype.is_complex
    T = P.random_peps(shape, D, seed=7, dtype=np.complex128 if cplx else np.float64)
    plan = AmpGradPlan(shape, D, chi, dtype=dtype)
    params = torch.tensor(_flat(T), dtype=dtype, device="cuda")
    n = shape[0] * shape[1]
    cfgs = [torch.tensor(P.random_configs(n, B, seed=s), dtype=torch.int32, device="cuda") for s in (1, 2, 3)]
    g = AmpGradGraph(plan, params, cfgs[0], want_p=True)
    for rep in range(3):
        for c in cfgs:
            g.configs.copy_(c)
            la, o, p = g.replay()
            la, o, p = la.clone(), o.clone(), p.clone()
            # eager work between replays, on another plan and on this one
            ref_la, ref_o, ref_p = plan.amp_grad(params, c, want_p=True)
            torch.cuda.synchronize()
            assert torch.equal(la, ref_la) and torch.equal(o, ref_o) and torch.equal(p, ref_p), (rep,)
    # new parameters copied into the graph's buffer are used by the next replay
    g.params.mul_(1.5)
    la, _, _ = g.replay()
    torch.cuda.synchronize()
    ref_la, _, _ = plan.amp_grad(params * 1.5, cfgs[-1])
    assert torch.equal(la, ref_la)


def test_mixed_sizes_one_graph_launch():
    """Standard and gauge-invariant plans of different sizes in one graph
    launch: every member's replay equals its eager evaluation."""
    from oracle import gi as G
    from vmc_amd import _lib
    from vmc_amd.engine import AmpGradPlan, GIPlan, MixedAmpGradGraph

    print("[graph] mixed sizes in one launch", flush=True)
    members, eager = [], []
    for shape, D, chi, B, seed in (((4, 4), 2, 4, 64, 1), ((6, 6), 3, 9, 32, 2), ((10, 10), 4, 32, 16, 3)):
        T = P.random_peps(shape, D, seed=seed)
        plan = AmpGradPlan(shape, D, chi, dtype=torch.float64)
        params = torch.tensor(_flat(T), device="cuda")
        cfg = torch.tensor(P.random_configs(shape[0] * shape[1], B, seed=seed), device="cuda")
        members.append((plan, params, cfg))
    for shape, dmax, chi, B, seed in (((4, 4), 2, 8, 32, 4), ((6, 6), 2, 8, 16, 5)):
        cfg = G.GIConfig(shape=shape, N=2, phys_dim=1, Qx=0, degeneracy_per_charge=(dmax, dmax), charge_of_site=(0,))
        T = G.random_gi_tensors(cfg, seed=seed)
        plan = GIPlan(shape, 2, 1, dmax, chi, (0,), strategy=_lib.VMC_VARIATIONAL, dtype=torch.float64)
        s = torch.tensor(G.random_configurations(cfg, B, seed=seed), device="cuda")
        members.append((plan, torch.tensor(_flat(T), device="cuda"), s))
    g = MixedAmpGradGraph(members)
    for plan, params, s in members:
        la, o, _ = plan.amp_grad(params, s)
        eager.append((la.clone(), o.clone()))
    for rep in range(3):
        outs = g.replay()
        torch.cuda.synchronize()
        for (la, o), (rla, ro) in zip(outs, eager):
            assert torch.equal(la, rla) and torch.equal(o, ro), rep
