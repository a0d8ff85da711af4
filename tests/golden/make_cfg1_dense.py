"""Writes cfg1_dense.npz: the Si2 FCC workload (a = 10.26 bohr, 15 Ha, 30^3 box, synthetic
norm-conserving species of seed 0) at three k-points of its 4x4x4 shifted Monkhorst-Pack list:
two low-symmetry ones (indices 6 and 3, no degeneracy among the lowest 9 levels) and the
(1/8, 1/8, 1/8) point (index 0), where the highest occupied level (band 4 of 4) is degenerate
with the lowest empty one.

Per k-point it stores the plane-wave list (g_index into the |G|-sorted density sphere, kinetic
|k+G|^2 in Ry, box offset), the 8 lowest eigenvalues of the dense Hamiltonian (Ry) and, for the
first k-point, a seeded two-band block psi with H psi.  The dense Hamiltonian is built
independently of any H-apply code: T on the diagonal, V(G - G') read from numpy's fftn of V(r)
divided by N (the FFTW_BACKWARD/FFTW_FORWARD pair of serial.hpp:456-478 reduces to this
convolution), plus beta D beta^H; scipy.linalg.eigh diagonalises it.  The plane-wave lists and
beta come from the oracle's setup, whose own pins are brute-force enumeration (set equality and
sort order) and the Gaussian closed form of the projector transform.

Run from the repository root: python tests/golden/make_cfg1_dense.py
"""
import os
import sys

import numpy as np
import scipy.linalg

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

from oracle import oracle as O  # noqa: E402
from q3_amd import workloads as W  # noqa: E402

KINDICES = (6, 3, 0)
NEIG = 8
PSI_BANDS = 2
PSI_SEED = 11


def dense_h(wave_dims, gi, ek, m, V, beta, bs, db):
    d = np.array(wave_dims)
    Vg = np.fft.fftn(V) / V.size
    mm = m[gi]
    w = np.mod(mm[:, None, :] - mm[None, :, :], d)
    H = Vg[w[..., 0], w[..., 1], w[..., 2]].astype(complex)
    H += np.diag(ek)
    offs = np.cumsum([0] + [x * x for x in bs[:-1]])
    D = scipy.linalg.block_diag(*[db[o:o + s * s].reshape(s, s, order="F") for o, s in zip(offs, bs)])
    return H + beta @ D @ beta.conj().T


def main():
    work = W.cfg1()
    cry = work.crystal
    b = O.Basis(cry.lattice, cry.alat, work.ecutwfc_ry, 0.0, True, work.fft_factors)
    m, g, g2 = b.gvectors()

    class _View:  # the basis attributes workloads.local_potential reads
        def g_vectors(self):
            return m, g, g2
        wave_dims = b.wave_dims
        cell_volume = abs(np.linalg.det(cry.lattice)) * cry.alat ** 3

    V = W.local_potential(cry, _View(), work.species)
    bs, db = W.d_blocks_for(cry, work.species)
    out = {"wave_dims": np.array(b.wave_dims, np.int32), "kindices": np.array(KINDICES, np.int32),
           "kpoints": work.kpoints[list(KINDICES)], "ecutwfc_ry": np.float64(work.ecutwfc_ry)}
    for j, ki in enumerate(KINDICES):
        k = work.kpoints[ki]
        gi, ek, fm = b.kpoint(k)
        beta = b.kb_projectors(k, gi, cry.positions, cry.species, work.species)
        H = dense_h(b.wave_dims, gi, ek, m, V, beta, bs, db)
        H = 0.5 * (H + H.conj().T)
        out[f"k{j}_g_index"] = gi.astype(np.int32)
        out[f"k{j}_kinetic_ry"] = ek
        out[f"k{j}_fft_map"] = fm.astype(np.int32)
        out[f"k{j}_eigenvalues_ry"] = scipy.linalg.eigh(H, eigvals_only=True, subset_by_index=[0, NEIG - 1])
        if j == 0:
            psi = W.random_psi(ek, PSI_BANDS, PSI_SEED)
            out["k0_psi"] = psi
            out["k0_hpsi"] = H @ psi
    np.savez_compressed(os.path.join(os.path.dirname(os.path.abspath(__file__)), "cfg1_dense.npz"), **out)


if __name__ == "__main__":
    main()
