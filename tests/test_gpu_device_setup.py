"""Device-side k-point setup: the projector set built on the device
(q3x_hamiltonian_build_projectors, the GPUProjectorSetBuilder of
hamiltonian_applicator/projector_set_builder_gpu.hpp) gives the H|psi> of the host-built set
(q3x_kb_projectors + set_projectors, itself <= 1e-13 from the oracle's projectors) to 1e-13
relative, for one and two species, l = 0..2, general k, the Gamma trick and the 64-atom cell; and
the device starting block equals the host generator's bit for bit."""
import numpy as np
import pytest

from q3_amd import FP64, PlanewaveHamiltonianApplicator
from q3_amd import workloads as W
from q3_amd.planewave import DeviceArray, PlanewaveBasis
from tests.helpers import rel_err

pytestmark = pytest.mark.gpu


def _spd_species(seed):
    """The synthetic species with a d channel added (l = 0, 1, 2, two betas of l = 1)."""
    sp = W.synthetic_species(seed)
    r = sp["r"]
    extra = [r * r * np.exp(-r * r / (2 * 0.55 ** 2)), r * r * r * np.exp(-r * r / (2 * 0.5 ** 2))]
    sp = dict(sp)
    sp["beta_r"] = np.vstack([sp["beta_r"], extra[0], extra[1]])
    sp["l"] = list(sp["l"]) + [1, 2]
    sp["cutoff"] = list(sp["cutoff"]) + [sp["cutoff"][0]] * 2
    n = len(sp["l"])
    a = np.random.default_rng(seed).uniform(1, 3, (n, n))
    sp["dion_ry"] = a + a.T
    return sp


CASES = {
    "cfg1_k5": (W.cfg1, 5, False, None),
    "cfg4_k1": (W.cfg4, 1, False, None),
    "cfg2_gamma_full": (W.cfg2, 0, False, None),
    "cfg1_gamma_trick": (W.cfg1, None, True, None),
    "alloy_two_species_spd": (None, None, False, "alloy"),
}


def _case(name):
    make, kidx, gamma, special = CASES[name]
    if special == "alloy":
        rng = np.random.default_rng(5)
        crystal = W.alloy_cell((1, 1, 1), "sc", 0.4, rng)
        species = [_spd_species(0), _spd_species(1)]
        b = PlanewaveBasis(crystal, 24.0, 0.0, True, (1, 1, 1))
        kp = b.compute_gk([0.13, -0.27, 0.31])
        return crystal, species, b, kp
    w = make()
    b = PlanewaveBasis(w.crystal, w.ecutwfc_ry, 0.0, True, w.fft_factors, gamma_only=gamma)
    kp = b.compute_gk([0, 0, 0] if gamma else w.kpoints[kidx])
    return w.crystal, w.species, b, kp


@pytest.mark.parametrize("name", list(CASES))
def test_device_projectors_equal_host_projectors(gpu_context, name):
    crystal, species, b, kp = _case(name)
    V = W.local_potential(crystal, b, species)
    bs, db = W.d_blocks_for(crystal, species)
    beta = kp.kb_projectors(crystal, species)
    psi = W.random_psi(kp.kinetic_energy_squared, 6, seed=3)
    if b.gamma_only:
        psi[0, :] = psi[0, :].real
    print(f"{name}: npw {kp.npw}, nproj {beta.shape[1]}", flush=True)
    h_host = PlanewaveHamiltonianApplicator(gpu_context, b)
    h_host.update_potential(V)
    h_host.prepare_kpoint(kp)
    h_host.set_projectors(beta, bs, db)
    h_dev = PlanewaveHamiltonianApplicator(gpu_context, b)
    h_dev.update_potential(V)
    h_dev.prepare_kpoint(kp)
    h_dev.build_projectors(crystal, species, bs, db)
    want = h_host.apply_batch(psi, FP64)
    got = h_dev.apply_batch(psi, FP64)
    assert rel_err(got, want) <= 1e-13
    # the nonlocal part alone (local parts are identical): H - H_local
    h_loc = PlanewaveHamiltonianApplicator(gpu_context, b)
    h_loc.update_potential(V)
    h_loc.prepare_kpoint(kp)
    loc = h_loc.apply_batch(psi, FP64)
    assert rel_err(got - loc, want - loc) <= 1e-12


def test_device_random_block_is_the_host_block(gpu_context):
    for make, gamma in ((W.cfg4, False), (W.cfg1, True)):
        w = make()
        b = PlanewaveBasis(w.crystal, w.ecutwfc_ry, 0.0, True, w.fft_factors, gamma_only=gamma)
        kp = b.compute_gk([0, 0, 0] if gamma else w.kpoints[2])
        h = PlanewaveHamiltonianApplicator(gpu_context, b)
        h.update_potential(W.local_potential(w.crystal, b, w.species))
        h.prepare_kpoint(kp)
        nb = 13
        d = DeviceArray(gpu_context, kp.npw * nb * 16)
        h.random_psi_device(nb, 77, d.ptr)
        gpu_context.synchronize()
        got = d.download((kp.npw, nb))
        want = kp.random_psi(nb, 77)
        assert got.tobytes(order="F") == want.tobytes(order="F")
