"""The SCF loop's GPU parts and the loop itself.

* Density accumulation on the device (q3x_hamiltonian_accumulate_density) against the oracle's
  numpy transform (oracle/scf_oracle.py:density, pinned by a closed form in test_scf_host.py):
  a cfg1 k-point with random bands and weights, and the Gamma trick with an odd band count
  (packed pairs and a lone last band), to 1e-12 relative; the cfg2 box (96^3, 160 bands) by
  Parseval (the box mean of sum w |psi|^2 is sum w sum |c|^2).
* The SCF loop (q3_amd.scf.PlanewaveSCF: Davidson and the density on the GPU) on cfg1 against the
  oracle's self-consistent solution (tests/golden/cfg1_scf.npz, tests/golden/make_cfg1_scf.py):
  total energy within 1e-8 Ha, every eigenvalue within 1e-6 Ha, Fermi level within 1e-6 Ha.
"""
import os

import numpy as np
import pytest

from oracle import scf_oracle as S
from q3_amd import PlanewaveHamiltonianApplicator
from q3_amd import workloads as W
from q3_amd.planewave import PlanewaveBasis
from q3_amd.scf import PlanewaveSCF

pytestmark = pytest.mark.gpu

HA = 2.0
FIX = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "cfg1_scf.npz")


def _rel(a, b):
    return float(np.abs(a - b).max() / np.abs(b).max())


def test_density_matches_the_oracle(gpu_context):
    w = W.cfg1()
    b = PlanewaveBasis(w.crystal, w.ecutwfc_ry, 0.0, True, w.fft_factors)
    kp = b.compute_gk(w.kpoints[5])
    h = PlanewaveHamiltonianApplicator(gpu_context, b)
    h.prepare_kpoint(kp)
    rng = np.random.default_rng(4)
    psi = rng.standard_normal((kp.npw, 8)) + 1j * rng.standard_normal((kp.npw, 8))
    wts = rng.uniform(0.1, 1.0, 8)
    got = h.density(psi, wts)
    want = S.density(b.wave_dims, kp.fft_map, psi, wts)
    print(f"cfg1 k-point density: {_rel(got, want):.1e} relative", flush=True)
    assert _rel(got, want) <= 1e-12


def test_gamma_trick_density_matches_the_oracle(gpu_context):
    w = W.cfg1()
    b = PlanewaveBasis(w.crystal, w.ecutwfc_ry, 0.0, True, w.fft_factors, gamma_only=True)
    kp = b.compute_gk([0.0, 0.0, 0.0])
    h = PlanewaveHamiltonianApplicator(gpu_context, b)
    h.prepare_kpoint(kp)
    psi = kp.random_psi(7, 11)  # real functions: G = 0 real
    wts = np.linspace(0.2, 1.0, 7)
    got = h.density(psi, wts)
    want = S.density(b.wave_dims, kp.fft_map, psi, wts, minus_map=kp.minus_fft_map)
    print(f"Gamma-trick density, 7 bands: {_rel(got, want):.1e} relative", flush=True)
    assert _rel(got, want) <= 1e-12


def test_cfg2_density_parseval(gpu_context):
    w = W.cfg2()
    b = PlanewaveBasis(w.crystal, w.ecutwfc_ry, 0.0, True, w.fft_factors)
    kp = b.compute_gk(w.kpoints[0])
    h = PlanewaveHamiltonianApplicator(gpu_context, b)
    h.prepare_kpoint(kp)
    psi = W.random_psi(kp.kinetic_energy_squared, 160, seed=2)
    wts = np.linspace(0.5, 1.5, 160)
    rho = h.density(psi, wts)
    want = float(np.sum(np.sum(np.abs(psi) ** 2, axis=0) * wts))
    print(f"cfg2 density mean {rho.mean():.15e}, Parseval {want:.15e}", flush=True)
    assert rho.min() >= 0 and abs(rho.mean() - want) <= 1e-12 * want


def test_scf_cfg1_against_the_oracle_solution(gpu_context):
    fx = np.load(FIX)
    w = W.cfg1()
    scf = PlanewaveSCF(gpu_context, w.crystal, w.species, w.ecutwfc_ry, w.kpoints, np.full(64, 1 / 64), w.nbands, 8,
                       fft_factors=w.fft_factors, sigma=float(fx["sigma"]), conv_thr=float(fx["conv_thr"]))
    try:
        res = scf.run(log=lambda m: print(m, flush=True))
    finally:
        scf.close()
    assert res.converged
    de = abs(res.etot - float(fx["etot_ry"])) / HA
    deig = float(np.abs(res.eigenvalues - fx["eigenvalues_ry"]).max()) / HA
    print(f"SCF: {res.iterations} iterations, E = {res.etot:.12f} Ry, |dE| = {de:.2e} Ha, max |d eig| = {deig:.2e} Ha",
          flush=True)
    assert de <= 1e-8
    assert deig <= 1e-6
    assert abs(res.fermi - float(fx["fermi_ry"])) / HA <= 1e-6
