"""Local operator terms (mirror of src/vmc/operators/local_terms.py:22-200).

Same term classes and ``bucket_terms`` grouping; ``to_tables`` flattens a
LocalHamiltonian into the arrays of the C entry point ``vmc_plan_set_terms``.
Operators are dense <out|op|in> matrices; the engine applies op[:, s] like the
reference (src/vmc/peps/common/energy.py:88, 179, 200).
"""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np


class LocalTerm:
    """Base class of local operator terms."""


@dataclass(frozen=True)
class OneSiteTerm(LocalTerm):
    row: int
    col: int
    op: np.ndarray

    @property
    def sites(self):
        return ((self.row, self.col),)


@dataclass(frozen=True)
class DiagonalTerm(LocalTerm):
    sites: tuple
    diag: np.ndarray


@dataclass(frozen=True)
class HorizontalTwoSiteTerm(LocalTerm):
    row: int
    col: int
    op: np.ndarray

    @property
    def sites(self):
        return ((self.row, self.col), (self.row, self.col + 1))


@dataclass(frozen=True)
class VerticalTwoSiteTerm(LocalTerm):
    row: int
    col: int
    op: np.ndarray

    @property
    def sites(self):
        return ((self.row, self.col), (self.row + 1, self.col))


@dataclass(frozen=True)
class PlaquetteTerm(LocalTerm):
    row: int
    col: int
    coeff: complex


@dataclass(frozen=True)
class LocalHamiltonian:
    shape: tuple
    terms: tuple = ()


def bucket_terms(terms, shape):
    """local_terms.py:165-200: group by type and lattice location."""
    n_rows, n_cols = shape
    one = [[[] for _ in range(n_cols)] for _ in range(n_rows)]
    hor = [[[] for _ in range(max(n_cols - 1, 0))] for _ in range(n_rows)]
    ver = [[[] for _ in range(n_cols)] for _ in range(max(n_rows - 1, 0))]
    plaq = [[[] for _ in range(max(n_cols - 1, 0))] for _ in range(max(n_rows - 1, 0))]
    diag = []
    for term in terms:
        if isinstance(term, OneSiteTerm):
            one[term.row][term.col].append(term)
        elif isinstance(term, HorizontalTwoSiteTerm):
            hor[term.row][term.col].append(term)
        elif isinstance(term, VerticalTwoSiteTerm):
            ver[term.row][term.col].append(term)
        elif isinstance(term, DiagonalTerm):
            diag.append(term)
        elif isinstance(term, PlaquetteTerm):
            plaq[term.row][term.col].append(term)
        else:
            raise TypeError(f"Unsupported term type: {type(term)!r}")
    return diag, one, hor, ver, plaq


def to_tables(operator: LocalHamiltonian, phys_dim: int):
    """Flatten terms into the device tables. The value arrays are float64, or
    complex128 when any term is complex (vmc_plan_set_terms_c)."""
    n_rows, n_cols = operator.shape
    d = phys_dim
    t = {k: [] for k in ("diag_sites", "diag_vals", "one_sites", "one_ops", "h_sites", "h_ops",
                         "v_sites", "v_ops")}
    for term in operator.terms:
        if isinstance(term, PlaquetteTerm):
            raise NotImplementedError("PlaquetteTerm applies to gauge-invariant PEPS, not the standard engine")
        if isinstance(term, DiagonalTerm):
            sites = [r * n_cols + c for r, c in term.sites]
            if len(sites) > 2:
                raise NotImplementedError("diagonal terms on more than two sites")
            dv = np.real_if_close(np.asarray(term.diag))
            vals = np.zeros(d * d, dtype=np.complex128 if np.iscomplexobj(dv) else np.float64)
            vals[: dv.size] = dv
            t["diag_sites"] += [sites[0], sites[1] if len(sites) == 2 else -1]
            t["diag_vals"].append(vals)
            continue
        op = np.real_if_close(np.asarray(term.op))
        flat = term.row * n_cols + term.col
        key = {OneSiteTerm: "one", HorizontalTwoSiteTerm: "h", VerticalTwoSiteTerm: "v"}[type(term)]
        t[f"{key}_sites"].append(flat)
        t[f"{key}_ops"].append(np.asarray(op, dtype=np.complex128 if np.iscomplexobj(op) else np.float64).reshape(-1))
    cplx = any(np.iscomplexobj(a) for k, v in t.items() if not k.endswith("sites") for a in v)
    vdt = np.complex128 if cplx else np.float64
    out = {}
    for k, v in t.items():
        if k.endswith("sites"):
            out[k] = np.asarray(v, dtype=np.int32)
        else:
            out[k] = np.concatenate(v).astype(vdt) if v else np.zeros(0, dtype=vdt)
    return out
