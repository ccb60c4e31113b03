"""A designed PEPS whose truncated boundary amplitudes are known in closed form.

Test infrastructure only (imported by tests/ and nothing in the product path).

Lattice 2 x 3, physical dimension 2, bond dimension D = 8. Row 0 routes a
"channel" index through its horizontal bonds so that, in the top sweep, every
theta matrix of the row is monomial (at most one nonzero per row and column):

* site (0, 0): theta_0[q0, j] = a0[j] if q0 == j < n0, else 0;
* site (0, 1): theta_1[(l, q1), j'] = carry[l, j] a1[j, q1] at j' = 2 j + q1, q1 < 2;
* site (0, 2): the last column, weights a2[j', q2] for q2 < 2 (no cut).

A monomial matrix's singular values are the moduli of its entries and its
singular vectors are unit vectors, so each cut keeps the chi largest entries:

* zip-up (src/vmc/peps/common/strategy.py:128-150) carries S V^H to the next
  column, so cut 1 ranks the paths (q0, q1) by the accumulated weight
  |a0[q0] a1[q0, q1]|;
* density-matrix (strategy.py:153-176) carries the isometry V^H, so cut 1
  ranks them by the current weight |a1[q0, q1]| alone.

Both keep the chi largest |a0| at cut 0. Row 1 selects one physical index per
column of the boundary (site tensors are unit vectors on the up bond with
trivial horizontal bonds), so the amplitude of a configuration is the weight
a0[q0] a1[q0, q1] a2[2 q0 + q1, q2] of the selected path if the path survived
both cuts, and 0 otherwise. With chi <= n0 every cut has at least chi nonzero
singular values, so neither strategy keeps null-space vectors and both results
are basis independent. Placing S on the wrong side of a zip-up cut turns one
rule into the other, which these amplitudes detect.
"""
from __future__ import annotations

import numpy as np

SHAPE = (2, 3)
D = 8
N0 = 4


def monomial_peps(seed, q0_pick="top"):
    """Site tensors (d, up, down, left, right) and the weights that define them.

    q0_pick: the two values of q0 selected by configurations with s(1, 0) = 0, 1;
    "top" picks the two largest a0, whose four paths compete at cut 1.
    """
    rng = np.random.default_rng(seed)
    a0 = np.exp(rng.uniform(np.log(0.05), np.log(20.0), N0))  # spread: the two keep rules often differ
    a1 = rng.uniform(0.3, 3.0, (N0, 2))
    a2 = rng.uniform(0.3, 3.0, (2 * N0, 2))
    if isinstance(q0_pick, str):
        q0_pick = tuple(int(i) for i in np.argsort(-a0)[:2])
    d = 2
    t00 = np.zeros((d, 1, D, 1, D))
    for j in range(N0):
        t00[:, 0, j, 0, j] = a0[j]
    t01 = np.zeros((d, 1, D, D, D))
    for j in range(N0):
        for q in range(2):
            t01[:, 0, q, j, 2 * j + q] = a1[j, q]
    t02 = np.zeros((d, 1, D, D, 1))
    for j in range(2 * N0):
        for q in range(2):
            t02[:, 0, q, j, 0] = a2[j, q]
    pick = [tuple(q0_pick), (0, 1), (0, 1)]
    row1 = []
    for c, (l, r) in enumerate([(1, D), (D, D), (D, 1)]):
        t = np.zeros((d, D, 1, l, r))
        for s in range(d):
            t[s, pick[c][s], 0, 0, 0] = 1.0
        row1.append(t)
    tensors = [[t00, t01, t02], row1]
    return tensors, (a0, a1, a2, pick)


def _top(values, chi):
    order = sorted(values, key=lambda kv: -abs(kv[1]))
    return {key for key, _ in order[:chi]}


def survivors(weights, chi, rule):
    """Paths (q0, q1) kept by both cuts under rule "zipup" or "density"."""
    a0, a1, _, _ = weights
    s0 = _top([(q0, a0[q0]) for q0 in range(N0)], chi)
    if rule == "zipup":
        cand = [((q0, q1), a0[q0] * a1[q0, q1]) for q0 in s0 for q1 in range(2)]
    elif rule == "density":
        cand = [((q0, q1), a1[q0, q1]) for q0 in s0 for q1 in range(2)]
    else:
        raise ValueError(rule)
    return _top(cand, chi)


def amplitudes(weights, chi, rule, configs):
    """Closed-form amplitudes of (z, 6) occupancy configurations."""
    a0, a1, a2, pick = weights
    keep = survivors(weights, chi, rule)
    out = np.zeros(len(configs))
    for i, cfg in enumerate(np.asarray(configs)):
        q0, q1, q2 = pick[0][cfg[3]], pick[1][cfg[4]], pick[2][cfg[5]]
        if (q0, q1) in keep:
            out[i] = a0[q0] * a1[q0, q1] * a2[2 * q0 + q1, q2]
    return out


def all_configs():
    return np.array([[(i >> (5 - b)) & 1 for b in range(6)] for i in range(64)], dtype=np.int32)
