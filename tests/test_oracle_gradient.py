"""Pins the oracle's env-gradients (src/vmc/peps/common/energy.py:259-301).

* central finite differences of the exact amplitude (recipe of
  tests/test_gradients_finite_diff.py:18-78);
* Euler identity: psi is linear in each site tensor, so
  sum_{udlr} A[s] * dpsi/dA[s] == psi at every site;
* the full-gradient layout scatters the sliced gradient into the active
  physical slice (recipe of tests/test_sequential_sampling.py:19-46).
"""
import numpy as np

from oracle import peps as P


def test_gradient_matches_central_differences():
    shape, D = (2, 3), 3
    T = P.random_peps(shape, D, seed=1)
    cfg = P.random_configs(6, 1, seed=1)
    strat = P.NoTruncation()
    _, grad, _ = P.value_and_grad(T, cfg, shape, strat, D, full_gradient=True)
    flat = np.concatenate([t.ravel() for row in T for t in row])
    shapes = [t.shape for row in T for t in row]

    def amp_of(v):
        out, off = [], 0
        for r in range(shape[0]):
            row = []
            for c in range(shape[1]):
                s = shapes[r * shape[1] + c]
                n = int(np.prod(s))
                row.append(v[off:off + n].reshape(s))
                off += n
            out.append(row)
        return P.value(out, cfg, shape, strat)[0]

    eps = 1e-6 * np.max(np.abs(flat))
    fd = np.zeros_like(flat)
    for i in range(flat.size):
        d = np.zeros_like(flat)
        d[i] = eps
        fd[i] = (amp_of(flat + d) - amp_of(flat - d)) / (2 * eps)
    assert np.max(np.abs(grad[0] - fd)) / np.max(np.abs(fd)) < 1e-6


def test_euler_identity_every_site():
    shape, D = (3, 3), 3
    T = P.random_peps(shape, D, seed=2)
    cfgs = P.random_configs(9, 8, seed=2)
    amp, grad, _ = P.value_and_grad(T, cfgs, shape, P.NoTruncation(), D)
    off = 0
    for r in range(3):
        for c in range(3):
            n = int(np.prod(T[r][c].shape[1:]))
            g = grad[:, off:off + n]
            A = T[r][c][cfgs[:, r * 3 + c]].reshape(len(cfgs), n)
            assert np.allclose(np.sum(A * g, axis=1), amp, rtol=1e-10, atol=0)
            off += n


def test_full_vs_sliced_layout():
    shape, D = (2, 3), 2
    T = P.random_peps(shape, D, seed=3)
    cfgs = P.random_configs(6, 5, seed=3)
    _, gs, p = P.value_and_grad(T, cfgs, shape, P.ZipUp(4), D)
    _, gf, pf = P.value_and_grad(T, cfgs, shape, P.ZipUp(4), D, full_gradient=True)
    assert pf is None
    assert gf.shape[1] == 2 * gs.shape[1]
    offs, offf = 0, 0
    for r in range(2):
        for c in range(3):
            n = int(np.prod(T[r][c].shape[1:]))
            full = gf[:, offf:offf + 2 * n].reshape(-1, 2, n)
            sel = full[np.arange(len(cfgs)), cfgs[:, r * 3 + c]]
            assert np.array_equal(sel, gs[:, offs:offs + n])
            other = full[np.arange(len(cfgs)), 1 - cfgs[:, r * 3 + c]]
            assert np.all(other == 0)
            assert np.all(p[:, offs:offs + n] == cfgs[:, [r * 3 + c]])
            offs += n
            offf += 2 * n
