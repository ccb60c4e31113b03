from vmc_amd.operators.local_terms import (
    DiagonalTerm,
    HorizontalTwoSiteTerm,
    LocalHamiltonian,
    LocalTerm,
    OneSiteTerm,
    PlaquetteTerm,
    VerticalTwoSiteTerm,
    bucket_terms,
    to_tables,
)

__all__ = [
    "DiagonalTerm", "HorizontalTwoSiteTerm", "LocalHamiltonian", "LocalTerm", "OneSiteTerm",
    "PlaquetteTerm", "VerticalTwoSiteTerm", "bucket_terms", "to_tables",
]
