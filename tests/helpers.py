"""Shared builders for the parity tests: one k-point of a workload, its projectors, D blocks
and V(r), produced by the product's setup (the oracle then consumes the same arrays)."""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np

from q3_amd import workloads as W
from q3_amd.planewave import KPoint, PlanewaveBasis


@dataclass
class KCase:
    w: W.Workload
    basis: PlanewaveBasis
    kp: KPoint
    V: np.ndarray
    beta: np.ndarray
    bsizes: np.ndarray
    dblocks: np.ndarray


_cache = {}


def kcase(work: W.Workload, kindex: int = 0) -> KCase:
    key = (work.name, work.ecutwfc_ry, kindex)
    if key in _cache:
        return _cache[key]
    b = PlanewaveBasis(work.crystal, work.ecutwfc_ry, 0.0, True, work.fft_factors)
    kp = b.compute_gk(work.kpoints[kindex])
    V = W.local_potential(work.crystal, b, work.species)
    beta = kp.kb_projectors(work.crystal, work.species)
    bs, db = W.d_blocks_for(work.crystal, work.species)
    c = KCase(work, b, kp, V, beta, bs, db)
    _cache[key] = c
    return c


def rel_err(a, b):
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))
