"""Bond-gauge directions lie in the null space of the device Jacobian.

A bond gauge A1 -> A1 X, A2 -> X^-1 A2 on the bond shared by two sites leaves
every exactly contracted amplitude unchanged, so to first order in X = 1 + eps K
the parameter direction v = (A1 K, -K A2) satisfies dpsi/dA . v = 0 for every
sample, and the QGT annihilates v (S v = 0). This fixes the layout of the
device's env-gradient, index by index: a transposed bond, a swapped up/down or
left/right block, or a gradient landing on the wrong site breaks it. Exact
contraction (NoTruncation) only: a truncating cut is not gauge invariant.
"""
import numpy as np
import pytest
import torch

from oracle import peps as P

pytestmark = pytest.mark.gpu


def _gauge_direction(T, shape, bond, K):
    """v for the bond ((r, c), (r2, c2)) with generator K, in the flat parameter order."""
    (r, c), (r2, c2) = bond
    dT = [[np.zeros_like(t) for t in row] for row in T]
    if r2 == r:  # horizontal: right index of (r, c), left index of (r, c + 1)
        dT[r][c] = np.einsum("sudlx,xy->sudly", T[r][c], K)
        dT[r2][c2] = -np.einsum("yx,sudxr->sudyr", K, T[r2][c2])
    else:  # vertical: down index of (r, c), up index of (r + 1, c)
        dT[r][c] = np.einsum("suxlr,xy->suylr", T[r][c], K)
        dT[r2][c2] = -np.einsum("yx,sxdlr->sydlr", K, T[r2][c2])
    return np.concatenate([t.ravel() for row in dT for t in row])


@pytest.mark.parametrize("dtype", [torch.float64, torch.complex128])
def test_bond_gauge_directions_are_in_the_jacobian_null_space(dtype):
    from vmc_amd.peps import PEPS, NoTruncation
    from vmc_amd.peps.compat import _value_and_grad

    shape, D = (3, 4), 3
    print(f"[gauge] bond-gauge null directions {shape} D={D} exact {dtype}", flush=True)
    npdt = np.complex128 if dtype.is_complex else np.float64
    T = P.random_peps(shape, D, seed=21, dtype=npdt)
    model = PEPS(shape=shape, bond_dim=D, contraction_strategy=NoTruncation(), tensors=T, dtype=dtype)
    cfg = torch.tensor(P.random_configs(shape[0] * shape[1], 48, seed=22))
    amp, grad, _ = _value_and_grad(model, cfg, full_gradient=True)
    g = grad.cpu().numpy()  # dpsi/dA, (B, n_params)
    rng = np.random.default_rng(23)
    bonds = [((0, 0), (0, 1)), ((1, 1), (1, 2)), ((2, 2), (2, 3)), ((0, 3), (1, 3)), ((1, 0), (2, 0)),
             ((1, 2), (2, 2))]
    for bond in bonds:
        K = rng.standard_normal((D, D))
        if dtype.is_complex:
            K = K + 1j * rng.standard_normal((D, D))
        v = _gauge_direction(T, shape, bond, K)
        dpsi = g @ v
        scale = np.abs(g) @ np.abs(v)  # the size of the terms that must cancel
        rel = np.max(np.abs(dpsi) / scale)
        print(f"  bond {bond}: max |dpsi . v| / sum |terms| = {rel:.1e}", flush=True)
        assert rel < 1e-12
        # the same K on one site only is not a gauge: the amplitude moves
        w = v.copy()
        off = 0
        for rr, row in enumerate(T):
            for cc, t in enumerate(row):
                if (rr, cc) == bond[1]:
                    w[off:off + t.size] = 0.0
                off += t.size
        assert np.max(np.abs(g @ w) / (np.abs(g) @ np.abs(w))) > 1e-3
    # S v = 0: the centred log-derivatives annihilate the direction, so does S = Oc^H Oc / Ns
    o = g / amp.cpu().numpy()[:, None]
    oc = o - o.mean(axis=0)
    v = _gauge_direction(T, shape, bonds[0], rng.standard_normal((D, D)))
    ocv = oc @ v
    assert np.linalg.norm(ocv) <= 1e-11 * np.linalg.norm(oc) * np.linalg.norm(v)
    sv = oc.conj().T @ ocv / o.shape[0]
    assert np.linalg.norm(sv) <= 1e-11 * np.linalg.norm(oc) ** 2 * np.linalg.norm(v) / o.shape[0]
