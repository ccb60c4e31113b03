"""Local terms of the gauge-invariant PEPS (mirror of src/vmc/peps/gi/local_terms.py)."""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np

from vmc_amd.operators.local_terms import (DiagonalTerm, HorizontalTwoSiteTerm, LocalTerm, OneSiteTerm,
                                           PlaquetteTerm, VerticalTwoSiteTerm)

__all__ = ["LinkDiagonalTerm", "MatterMassTerm", "GILocalHamiltonian", "build_electric_terms", "gi_tables"]


@dataclass(frozen=True)
class LinkDiagonalTerm(DiagonalTerm):
    """local_terms.py:13-36: diag[link value] summed over the listed links of
    one orientation ('h' or 'v')."""

    orientation: str = "h"

    def __post_init__(self):
        if self.orientation not in ("h", "v"):
            raise ValueError("orientation must be 'h' or 'v'")

    def energy(self, h_links, v_links):
        links = h_links if self.orientation == "h" else v_links
        return sum(np.asarray(self.diag)[links[r, c]] for r, c in self.sites)


class MatterMassTerm(DiagonalTerm):
    """local_terms.py:39-52: coeff * charge(sigma) on one matter site."""

    def __init__(self, *, row, col, coeff, charge_of_site):
        super().__init__(sites=((row, col),), diag=np.asarray(charge_of_site, dtype=np.int64) * coeff)


@dataclass(frozen=True)
class GILocalHamiltonian:
    """local_terms.py:55-71."""

    shape: tuple
    terms: tuple = ()


def build_electric_terms(shape, coeff, N):
    """local_terms.py:74-90: coeff * (2 - 2 cos(2 pi k / N)) on every link."""
    n_rows, n_cols = shape
    diag = coeff * (2.0 - 2.0 * np.cos(2.0 * np.pi * np.arange(N) / N))
    terms = [LinkDiagonalTerm(sites=((r, c),), diag=diag, orientation="h")
             for r in range(n_rows) for c in range(n_cols - 1)]
    terms += [LinkDiagonalTerm(sites=((r, c),), diag=diag, orientation="v")
              for r in range(n_rows - 1) for c in range(n_cols)]
    return tuple(terms)


def gi_tables(operator: GILocalHamiltonian, N: int, phys_dim: int = 1):
    """Device tables:
    * per-plaquette summed coefficients (NaN where no term, gi/model.py:577-582);
    * the (n_links, N) link-diagonal table;
    * the terms on matter states in the layout of ``vmc_plan_set_terms``, or
      None: DiagonalTerms (MatterMassTerm among them, gi/model.py:422-431) and
      one-site / horizontal / vertical operators with the links held fixed
      (gi/model.py:474-536)."""
    n_rows, n_cols = operator.shape
    nh = n_rows * (n_cols - 1)
    plaq = np.full(max((n_rows - 1) * (n_cols - 1), 0), np.nan)
    link = np.zeros((nh + (n_rows - 1) * n_cols, N))
    has_link = False
    dsites, dvals = [], []
    ops = {"one": ([], []), "h": ([], []), "v": ([], [])}
    d = phys_dim
    for t in operator.terms:
        if isinstance(t, PlaquetteTerm):
            i = t.row * (n_cols - 1) + t.col
            coeff = complex(t.coeff)
            if coeff.imag != 0:
                raise NotImplementedError("complex plaquette coefficients")
            plaq[i] = coeff.real if np.isnan(plaq[i]) else plaq[i] + coeff.real
        elif isinstance(t, LinkDiagonalTerm):
            has_link = True
            for r, c in t.sites:
                l = r * (n_cols - 1) + c if t.orientation == "h" else nh + r * n_cols + c
                link[l] += np.asarray(t.diag, dtype=np.float64)
        elif isinstance(t, DiagonalTerm):
            sites = [r * n_cols + c for r, c in t.sites]
            if len(sites) > 2:
                raise NotImplementedError("diagonal terms on more than two sites")
            dv = np.real_if_close(np.asarray(t.diag))
            if np.iscomplexobj(dv):
                raise NotImplementedError("complex diagonal terms")
            vals = np.zeros(d * d)
            vals[: dv.size] = dv
            dsites += [sites[0], sites[1] if len(sites) == 2 else -1]
            dvals.append(vals)
        elif isinstance(t, (OneSiteTerm, HorizontalTwoSiteTerm, VerticalTwoSiteTerm)):
            op = np.real_if_close(np.asarray(t.op))
            if np.iscomplexobj(op):
                raise NotImplementedError("complex operators on matter states")
            key = {OneSiteTerm: "one", HorizontalTwoSiteTerm: "h", VerticalTwoSiteTerm: "v"}[type(t)]
            ops[key][0].append(t.row * n_cols + t.col)
            ops[key][1].append(np.asarray(op, dtype=np.float64).reshape(-1))
        else:
            raise NotImplementedError(f"{type(t).__name__} is not supported by the device GI estimate")
    diag = None
    if dsites or any(ops[k][0] for k in ops):
        diag = {"diag_sites": np.asarray(dsites, np.int32),
                "diag_vals": np.concatenate(dvals) if dvals else np.zeros(0)}
        for k, (sites_k, ops_k) in ops.items():
            diag[f"{k}_sites"] = np.asarray(sites_k, np.int32)
            diag[f"{k}_ops"] = np.concatenate(ops_k) if ops_k else np.zeros(0)
    return (plaq if not np.all(np.isnan(plaq)) else None), (link if has_link else None), diag
