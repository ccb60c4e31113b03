"""Shape-replay counter of the algorithmic work of one amp+grad sample.

Replays the reference's contraction orders on shapes only:
* zip-up theta: ``carry . m`` then ``. w`` (src/vmc/peps/common/strategy.py:95-115);
* truncation SVD of the (left*q) x (Dr*wr) theta matrix (strategy.py:142-148),
  priced with the Golub-Van Loan R-SVD count 6 M N^2 + 20 N^3 (M >= N);
* right environments ``(top.R).mpo.bot`` (src/vmc/peps/common/contraction.py:68-86);
* env-gradient ``(L.top), (bot.R), .`` and left update ``(L.top).mpo.bot``
  (src/vmc/peps/common/energy.py:98-120, 259-273).

The top sweep applies all rows; the bottom sweep applies rows R-1..1 (the
reference also applies row 0 from below and discards it, energy.py:299; that
discarded step is not counted). Multiply-accumulates are real MACs; FLOP = 2 MAC.
"""
from __future__ import annotations

from dataclasses import dataclass, field


def site_dims(r, c, R, C, D):
    return (1 if r == 0 else D, 1 if r == R - 1 else D, 1 if c == 0 else D, 1 if c == C - 1 else D)


@dataclass
class Work:
    theta_mac: int = 0
    row_mac: int = 0
    svd_flop: int = 0          # R-SVD count of the reference's truncations
    trunc_exec_flop: int = 0   # FP64 flops the Gram/tridiagonal route executes
    svd_shapes: dict = field(default_factory=dict)
    trunc_elems: int = 0       # compulsory elements per truncation: theta read, U and carry written
    launches: dict = field(default_factory=dict)

    @property
    def contraction_flop(self):
        return 2 * (self.theta_mac + self.row_mac)


def _zipup(work, bnd, R, C, D, chi, row, from_below):
    out = []
    k_prev = None
    for c in range(C):
        u, dn, l, r = site_dims(row, c, R, C, D)
        P, Q = (dn, u) if from_below else (u, dn)
        Dl, Pin, Dr = bnd[c]
        assert Pin == P
        left = Dl * l if c == 0 else k_prev
        if c > 0:
            work.theta_mac += k_prev * l * Dl * P * Dr
        work.theta_mac += left * Dr * l * P * Q * r
        ncols = Dr * r
        if c == C - 1:
            out.append((left, Q, ncols))
            break
        m = left * Q
        k = min(chi, m, ncols)
        M, N = max(m, ncols), min(m, ncols)
        work.svd_flop += 6 * M * N * N + 20 * N ** 3
        work.trunc_exec_flop += (
            0 if k == m else 2 * m * m * ncols + (4 * m ** 3) // 3 + 2 * m * m * k + 2 * m * k * ncols
        )
        work.svd_shapes[(m, ncols)] = work.svd_shapes.get((m, ncols), 0) + 1
        work.trunc_elems += m * ncols + m * k + k * ncols
        out.append((left, Q, k))
        k_prev = k
    return out


def _row_envs(work, top, bot, R, C, D, row):
    mac = 0
    # right environments
    for c in range(C - 2, -1, -1):
        cc = c + 1
        ta, tu, tb = top[cc]
        be, bv, bf = bot[cc]
        _, _, wl, wr = site_dims(row, cc, R, C, D)
        mac += ta * tu * tb * wr * bf       # X1 = top . R
        mac += wl * bv * wr * tu * ta * bf  # X2 = mpo . X1
        mac += ta * wl * bv * bf * be       # R = X2 . bot
    # gradients and left environments
    for c in range(C):
        ta, tu, tb = top[c]
        be, bv, bf = bot[c]
        _, _, wl, wr = site_dims(row, c, R, C, D)
        mac += wl * be * ta * tu * tb       # Y1 = L . top
        mac += be * bv * bf * tb * wr       # Y2 = bot . R
        mac += wl * tu * be * tb * bv * wr  # grad = Y1 . Y2
        if c < C - 1:
            mac += wr * bv * wl * tu * be * tb  # Z2 = mpo . Y1
            mac += tb * wr * be * bv * bf       # L = Z2 . bot
    work.row_mac += mac


def amp_grad_work(shape, D, chi, grad=True):
    """Algorithmic work of one sample of _value_and_grad (compat.py:94-133)."""
    R, C = shape
    w = Work()
    tops = [[(1, 1, 1)] * C]
    for r in range(R):
        tops.append(_zipup(w, tops[-1], R, C, D, chi, r, False))
    w.launches = {"theta": R * C, "trunc": R * (C - 1)}
    if not grad:
        return w
    bot = [(1, 1, 1)] * C
    for r in range(R - 1, -1, -1):
        _row_envs(w, tops[r], bot, R, C, D, r)
        if r > 0:
            bot = _zipup(w, bot, R, C, D, chi, r, True)
    w.launches["theta"] += (R - 1) * C
    w.launches["trunc"] += (R - 1) * (C - 1)
    w.launches["row_grad"] = R
    return w
