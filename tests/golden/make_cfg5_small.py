"""Writes cfg5_small.npz: eigenvalues (Ry) of a small request of the serving stream, computed by
the FP64 CPU oracle, for the GPU server test at a size the oracle covers.

The request is the first 2-atom cell of the cfg5 stream (Poisson 20 req/s, seed 4: request 6, a
Si/C primitive cell), its cutoff and band count, and the first two k-points of its path.  Every
part comes from the oracle: G sphere and plane-wave lists, KB projectors and D blocks of the two
synthetic species, and its block Davidson with every band held to 1e-11 Ry from a damped random
start; V(r) is the workload's Gaussian superposition evaluated in numpy on the oracle's G sphere.
The residual norms ||H psi - e psi|| of the returned vectors (oracle H-apply) are stored beside
the eigenvalues.

Run from the repository root (seconds):
    python tests/golden/make_cfg5_small.py
"""
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

from oracle import oracle as O  # noqa: E402
from q3_amd import workloads as W  # noqa: E402

REQUEST = 6
NK = 2


class _OracleBasis:
    def __init__(self, ob, crystal):
        self.ob = ob
        self.wave_dims = ob.wave_dims
        self.gamma_only = False
        _r, self.cell_volume, _t = O.crystal(crystal.lattice, crystal.alat)

    def g_vectors(self):
        return self.ob.gvectors()


def main():
    r = W.cfg5_requests(REQUEST + 1, seed=4)[REQUEST]
    assert len(r.crystal.species) == 2
    c = r.crystal
    sp = W.serving_species()
    ob = O.Basis(c.lattice, c.alat, r.ecutwfc_ry, 0.0, True, (1, 1, 1))
    V = W.local_potential(c, _OracleBasis(ob, c), sp)
    bs, db = O.d_blocks(c.species, sp)
    eigs, res = [], []
    for i in range(NK):
        k = r.kpoints[i]
        gi, ek, fm = ob.kpoint(k)
        beta = ob.kb_projectors(k, gi, c.positions, c.species, sp)
        psi0 = W.random_psi(ek, r.nbands, seed=50 + i)
        e, psi, rep = O.davidson(ob.wave_dims, fm, ek, V, beta, bs, db, r.nbands, r.nbands, psi0=psi0, tol=1e-11,
                                 max_iter=500)
        assert rep["unconverged"] == 0
        h = O.hamiltonian_apply(ob.wave_dims, fm, ek, V, beta, bs, db, psi)
        eigs.append(e)
        res.append(np.linalg.norm(h - psi * e[None, :], axis=0))
        print(f"k {i}: npw {len(gi)}, {rep['iterations']} iterations, max residual {res[-1].max():.2e}")
    out = os.path.join(os.path.dirname(os.path.abspath(__file__)), "cfg5_small.npz")
    np.savez_compressed(out, request=REQUEST, eigenvalues_ry=np.array(eigs), residual_norms=np.array(res),
                        kpoints=np.asarray(r.kpoints[:NK]), ecutwfc_ry=r.ecutwfc_ry, nbands=r.nbands)


if __name__ == "__main__":
    main()
