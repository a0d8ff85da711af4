"""The eigenvalues_at_kpoints results record (the reference's properties schema,
src/generated/esse/properties/non_scalar/eigenvalues_at_kpoints.hpp): name, units and, per
k-point, its coordinates, weight and per-spin band energies and occupations, so band energies
from this path can be compared with the reference's tools.

    {"name": "eigenvalues_at_kpoints", "units": "Ry",
     "values": [{"kpoint": [kx, ky, kz], "weight": w,
                 "eigenvalues": [{"spin": 0.5, "energies": [...], "occupations": [...]}]}]}

The solver works in Ry; a record in hartree or eV converts on writing (1 Ry = 0.5 hartree =
13.605693122994 eV, CODATA 2018).  The units the schema names but that are per-mole or per-atom
(kcal/mol, kJ/mol, J/mol, cm-1, eV/atom) are not written; reading refuses a name the schema does
not list, as the schema's own reader does.
"""
from __future__ import annotations

import numpy as np

NAME = "eigenvalues_at_kpoints"
SCHEMA_UNITS = ("kcal/mol", "kJ/mol", "eV", "J/mol", "hartree", "cm-1", "Ry", "eV/atom")
_PER_RY = {"Ry": 1.0, "hartree": 0.5, "eV": 13.605693122994}


def eigenvalues_at_kpoints(kpoints, eigenvalues_ry, units: str = "Ry", weights=None, occupations=None,
                           spin=None) -> dict:
    """Record of nk x nb band energies (Ry in) at nk k-points (crystal coordinates)."""
    if units not in _PER_RY:
        raise ValueError(f"units {units!r}: written in one of {sorted(_PER_RY)}")
    k = np.asarray(kpoints, dtype=np.float64).reshape(-1, 3)
    e = np.asarray(eigenvalues_ry, dtype=np.float64).reshape(len(k), -1) * _PER_RY[units]
    occ = None if occupations is None else np.asarray(occupations, dtype=np.float64).reshape(e.shape)
    values = []
    for i in range(len(k)):
        band = {"energies": e[i].tolist()}
        if spin is not None:
            band["spin"] = float(spin)
        if occ is not None:
            band["occupations"] = occ[i].tolist()
        entry = {"kpoint": k[i].tolist(), "eigenvalues": [band]}
        if weights is not None:
            entry["weight"] = float(weights[i])
        values.append(entry)
    return {"name": NAME, "units": units, "values": values}


def read_eigenvalues_at_kpoints(record: dict):
    """(kpoints nk x 3, energies nk x nb in Ry, weights or None) of a record (first spin channel)."""
    if record.get("name") != NAME:
        raise ValueError(f"record name {record.get('name')!r} is not {NAME!r}")
    units = record.get("units")
    if units not in SCHEMA_UNITS:
        raise ValueError(f"unknown units {units!r}")
    if units not in _PER_RY:
        raise ValueError(f"units {units!r} are not an energy per electron")
    vals = record["values"]
    k = np.array([v.get("kpoint", [np.nan] * 3) for v in vals], dtype=np.float64).reshape(-1, 3)
    e = np.array([v["eigenvalues"][0]["energies"] for v in vals], dtype=np.float64) / _PER_RY[units]
    w = [v.get("weight") for v in vals]
    return k, e, (None if any(x is None for x in w) else np.array(w, dtype=np.float64))
