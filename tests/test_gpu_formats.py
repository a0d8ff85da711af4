"""The on-disk formats on the GPU path: band energies of cfg1 k-points computed on the device from
species read out of UPF files (the synthetic species written as UPF v2, and the reference's own
Si.pz-vbc.UPF v1 file, tests/golden/), written as the reference's eigenvalues_at_kpoints record
and read back; compared with the FP64 oracle run on the same files' species.

Bars: eigenvalues of the file species equal those of the in-memory species bit for bit (the file
holds the same numbers); every band within 1e-8 Ha of the oracle (every band held to 1e-10 Ry);
the record round trip is exact."""
import json
import os

import numpy as np
import pytest

from oracle import oracle as O
from q3_amd import PlanewaveDavidsonDiagonalizer, PlanewaveHamiltonianApplicator, records, upf
from q3_amd import workloads as W
from q3_amd.planewave import PlanewaveBasis

pytestmark = pytest.mark.gpu

HA = 2.0
GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def _solve(ctx, w, species, kidx):
    b = PlanewaveBasis(w.crystal, w.ecutwfc_ry, 0.0, True, w.fft_factors)
    kp = b.compute_gk(w.kpoints[kidx])
    h = PlanewaveHamiltonianApplicator(ctx, b)
    # the local part is the workloads' Gaussian form (Z = 4, r_loc = 0.44 bohr) for both files; the
    # files provide the projectors and D
    V = W.local_potential(w.crystal, b, [{"z_val": 4.0, "r_loc": 0.44}])
    h.update_potential(V)
    h.prepare_kpoint(kp)
    bs, db = W.d_blocks_for(w.crystal, species)
    h.build_projectors(w.crystal, species, bs, db)
    d = PlanewaveDavidsonDiagonalizer(h)
    d.set_tolerance(1e-10)
    d.set_maximum_iterations(300)
    e, _ = d.solve(w.nbands, w.nbands, kp.random_psi(w.nbands, 7))
    return e, b, kp, V, bs, db


@pytest.mark.parametrize("fname", ["si_synthetic_seed0.upf", "Si.pz-vbc.UPF"])
def test_band_energies_from_upf_files_through_the_record(gpu_context, tmp_path, fname):
    w = W.cfg1()
    sp_file = upf.read_upf(os.path.join(GOLDEN, fname))
    if fname.startswith("si_synthetic"):
        sp_mem = W.synthetic_species(0)
        e_mem, *_ = _solve(gpu_context, w, [sp_mem], 0)
    eigs, kpts = [], []
    for kidx in (0, 4):
        e, b, kp, V, bs, db = _solve(gpu_context, w, [sp_file], kidx)
        # the oracle on the same file species
        ob = O.Basis(w.crystal.lattice, w.crystal.alat, w.ecutwfc_ry, 0.0, True, w.fft_factors)
        gi, ek, fm = ob.kpoint(w.kpoints[kidx])
        beta = ob.kb_projectors(w.kpoints[kidx], gi, w.crystal.positions, w.crystal.species, [sp_file])
        obs, odb = O.d_blocks(w.crystal.species, [sp_file])
        oe, _, orep = O.davidson(ob.wave_dims, fm, ek, V, beta, obs, odb, w.nbands, w.nbands,
                                 W.random_psi(ek, w.nbands, seed=7), tol=1e-10, max_iter=300)
        assert orep["unconverged"] == 0
        assert np.abs(e - oe).max() <= 1e-8 * HA
        if kidx == 0 and fname.startswith("si_synthetic"):
            assert np.array_equal(e, e_mem)
        eigs.append(e)
        kpts.append(w.kpoints[kidx])
    rec = records.eigenvalues_at_kpoints(np.array(kpts), np.array(eigs), units="Ry")
    path = tmp_path / "eigenvalues.json"
    path.write_text(json.dumps(rec))
    k2, e2 = records.read_eigenvalues_at_kpoints(json.loads(path.read_text()))[:2]
    np.testing.assert_array_equal(np.asarray(e2), np.array(eigs))
    np.testing.assert_array_equal(np.asarray(k2), np.array(kpts))
