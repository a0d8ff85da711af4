"""GPU parity of H|psi> (FFT sandwich + kinetic + KB nonlocal) against the FP64 oracle.

Bars: FP64 relative error <= 1e-10 (norm over the block); mixed precision (complex64
FFT / projector GEMMs, fp32 accumulation) <= 2e-6.  Full-size properties at cfg2:
Hermiticity and linearity, which hold independently of any reference.
"""
import numpy as np
import pytest

from oracle import oracle as O
from q3_amd import FP64, MIXED, PlanewaveHamiltonianApplicator
from q3_amd import workloads as W
from q3_amd._lib import StateError
from tests.helpers import kcase, rel_err

pytestmark = pytest.mark.gpu


def _applicator(ctx, case, projectors=True):
    h = PlanewaveHamiltonianApplicator(ctx, case.basis)
    h.update_potential(case.V)
    h.prepare_kpoint(case.kp)
    if projectors:
        h.set_projectors(case.beta, case.bsizes, case.dblocks)
    return h


def _oracle(case, psi, projectors=True):
    kp = case.kp
    if projectors:
        return O.hamiltonian_apply(case.basis.wave_dims, kp.fft_map, kp.kinetic_energy_squared, case.V, case.beta,
                                   case.bsizes, case.dblocks, psi)
    return O.hamiltonian_apply(case.basis.wave_dims, kp.fft_map, kp.kinetic_energy_squared, case.V, None, None, None,
                               psi)


def test_apply_before_prepare_raises(gpu_context):
    case = kcase(W.cfg1(), 0)
    h = PlanewaveHamiltonianApplicator(gpu_context, case.basis)
    h.update_potential(case.V)
    with pytest.raises(StateError):
        h.apply_batch(np.zeros((case.kp.npw, 1), np.complex128))


@pytest.mark.parametrize("kindex", [0, 5, 37])
def test_local_part_fp64_cfg1(gpu_context, kindex):
    case = kcase(W.cfg1(), kindex)
    psi = W.random_psi(case.kp.kinetic_energy_squared, 8, seed=11 + kindex)
    h = _applicator(gpu_context, case, projectors=False)
    got = h.apply_batch(psi, FP64)
    ref = _oracle(case, psi, projectors=False)
    assert rel_err(got, ref) <= 1e-10


def test_ragged_and_empty_band_counts(gpu_context):
    """Band counts off every internal tile and batch size agree with the oracle column by
    column; nvec = 0 is a no-op and a negative count is refused."""
    from q3_amd._lib import Q3XError, lib
    case = kcase(W.cfg1(), 9)
    h = _applicator(gpu_context, case)
    wide = W.random_psi(case.kp.kinetic_energy_squared, 65, seed=21)
    ref = _oracle(case, wide)
    for nvec in (1, 2, 7, 33, 65):
        assert rel_err(h.apply_batch(wide[:, :nvec], FP64), ref[:, :nvec]) <= 1e-12
        assert rel_err(h.apply_batch(wide[:, :nvec], MIXED), ref[:, :nvec]) <= 2e-6
    assert h.apply_batch(np.zeros((case.kp.npw, 0), np.complex128)).shape == (case.kp.npw, 0)
    with pytest.raises(Q3XError):
        from q3_amd._lib import check
        check(lib().q3x_hamiltonian_apply(h.h, None, None, -1, FP64))


@pytest.mark.parametrize("kindex", [0, 21])
def test_full_h_fp64_cfg1(gpu_context, kindex):
    case = kcase(W.cfg1(), kindex)
    psi = W.random_psi(case.kp.kinetic_energy_squared, 8, seed=3 + kindex)
    h = _applicator(gpu_context, case)
    got = h.apply_batch(psi, FP64)
    ref = _oracle(case, psi)
    assert rel_err(got, ref) <= 1e-10


def test_full_h_mixed_cfg1(gpu_context):
    case = kcase(W.cfg1(), 3)
    psi = W.random_psi(case.kp.kinetic_energy_squared, 8, seed=5)
    h = _applicator(gpu_context, case)
    got = h.apply_batch(psi, MIXED)
    ref = _oracle(case, psi)
    assert rel_err(got, ref) <= 2e-6


def test_constant_potential_closed_form(gpu_context):
    """V(r) = v0 everywhere: H psi = (T + v0) psi exactly (no projectors)."""
    case = kcase(W.cfg1(), 1)
    v0 = -0.37
    h = PlanewaveHamiltonianApplicator(gpu_context, case.basis)
    h.update_potential(np.full(case.V.shape, v0))
    h.prepare_kpoint(case.kp)
    psi = W.random_psi(case.kp.kinetic_energy_squared, 5, seed=2)
    got = h.apply_batch(psi, FP64)
    want = (case.kp.kinetic_energy_squared + v0)[:, None] * psi
    assert rel_err(got, want) <= 1e-13


@pytest.mark.parametrize("m0", [(1, -2, 3), (0, 3, -1)])
def test_single_harmonic_potential_closed_form(gpu_context, m0):
    """V(r) = sin(G0 . r): H psi = T psi - i/2 psi(G - G0) + i/2 psi(G + G0) (cyclic on the
    box), computed without an FFT; FP64 to 1e-13, complex64 H|psi> to 2e-6."""
    from tests.helpers import sine_h_closed_form, sine_potential
    case = kcase(W.cfg1(), 1)
    b, kp = case.basis, case.kp
    h = PlanewaveHamiltonianApplicator(gpu_context, b)
    h.update_potential(sine_potential(case.V.shape, m0))
    h.prepare_kpoint(kp)
    psi = W.random_psi(kp.kinetic_energy_squared, 6, seed=5)
    miller = b.g_vectors()[0][kp.g_vector_indices]
    want = sine_h_closed_form(miller, kp.fft_map, b.wave_dims, m0, kp.kinetic_energy_squared, psi)
    assert rel_err(h.apply_batch(psi, FP64), want) <= 1e-13
    assert rel_err(h.apply_batch(psi, MIXED), want) <= 2e-6


@pytest.mark.parametrize("alat,ecut", [(8.0, 20.0), (9.3, 31.0), (6.1, 44.0)])
def test_triclinic_boxes_fp64(gpu_context, alat, ecut):
    """Unequal, non-cubic boxes (n1 != n2 != n3, radix plans with 2/3/5 stages)."""
    from q3_amd.planewave import Crystal
    lat = np.array([[1.0, 0.0, 0.0], [0.31, 1.07, 0.0], [0.22, 0.13, 1.29]])
    c = Crystal(lat, alat, np.array([[0.0, 0.0, 0.0], [0.27, 0.31, 0.23]]), np.zeros(2, np.int32))
    w = W.Workload("triclinic", c, ecut, (1, 1, 1), np.array([[0.1, -0.2, 0.3]]), 6, 3, [W.synthetic_species(0)])
    case = kcase(w, 0)
    assert len(set(case.basis.wave_dims)) == 3
    psi = W.random_psi(case.kp.kinetic_energy_squared, 6, seed=4)
    h = _applicator(gpu_context, case)
    got = h.apply_batch(psi, FP64)
    ref = _oracle(case, psi)
    assert rel_err(got, ref) <= 1e-10


def test_cfg4_kpoint_fp64(gpu_context):
    case = kcase(W.cfg4(), 7)
    psi = W.random_psi(case.kp.kinetic_energy_squared, 6, seed=9)
    h = _applicator(gpu_context, case)
    got = h.apply_batch(psi, FP64)
    ref = _oracle(case, psi)
    assert rel_err(got, ref) <= 1e-10


def test_cfg2_bands_fp64_and_mixed(gpu_context):
    case = kcase(W.cfg2(), 0)
    psi = W.random_psi(case.kp.kinetic_energy_squared, 8, seed=2)
    h = _applicator(gpu_context, case)
    ref = _oracle(case, psi)
    got = h.apply_batch(psi, FP64)
    assert rel_err(got, ref) <= 1e-10
    got32 = h.apply_batch(psi, MIXED)
    assert rel_err(got32, ref) <= 2e-6


def test_cfg2_full_block_hermitian_and_linear(gpu_context):
    """All 160 bands at cfg2: <phi|H psi> = conj(<psi|H phi>) and H(a x + b y) = a Hx + b Hy."""
    case = kcase(W.cfg2(), 0)
    npw = case.kp.npw
    X = W.random_psi(case.kp.kinetic_energy_squared, 160, seed=2)
    h = _applicator(gpu_context, case)
    HX = h.apply_batch(X, FP64)
    G = X.conj().T @ HX
    assert np.abs(G - G.conj().T).max() <= 1e-10 * np.abs(G).max()
    a, b = 0.3 - 1.1j, 2.0 + 0.5j
    Y = np.asfortranarray(a * X[:, :80] + b * X[:, 80:])
    HY = h.apply_batch(Y, FP64)
    assert rel_err(HY, a * HX[:, :80] + b * HX[:, 80:]) <= 1e-12
    assert HX.shape == (npw, 160)


# box sizes of the two-level plane and z-stick kernels (plane_tl.hip)
TL_SIZES = [24, 30, 32, 36, 40, 45, 48, 54, 60, 64, 72, 75, 80, 81, 90, 96, 100, 108, 120, 128, 135, 144]


def _cubic_case(n):
    """Simple-cubic one-atom cell whose wave grid is n^3: extent e = (n - 1) // 2 along each
    axis, ecutwfc chosen so that the 4 x ecutwfc sphere reaches e + 1/2 (in 2 pi / a units)."""
    from q3_amd.planewave import Crystal
    alat = 10.0
    e = (n - 1) // 2
    ecut = ((e + 0.5) * 2.0 * np.pi / alat) ** 2 / 4.0
    c = Crystal(np.eye(3), alat, np.array([[0.1, 0.2, 0.3]]), np.zeros(1, np.int32))
    w = W.Workload(f"sc{n}", c, ecut, (1, 1, 1), np.array([[0.15, -0.1, 0.05]]), 2, 1, [W.synthetic_species(0)])
    case = kcase(w, 0)
    assert case.basis.wave_dims == (n, n, n)
    return case


@pytest.mark.parametrize("n", TL_SIZES)
def test_two_level_plane_sizes(gpu_context, n):
    """Every box size with two-level plane and z-stick kernels, FP64 (1e-10) and mixed (2e-6),
    local part against the oracle."""
    case = _cubic_case(n)
    psi = W.random_psi(case.kp.kinetic_energy_squared, 2, seed=n)
    h = _applicator(gpu_context, case, projectors=False)
    ref = _oracle(case, psi, projectors=False)
    assert rel_err(h.apply_batch(psi, FP64), ref) <= 1e-10
    assert rel_err(h.apply_batch(psi, MIXED), ref) <= 2e-6


def test_fft_kernels_agree_cfg2(gpu_context, monkeypatch):
    """The generic Stockham kernels (Q3X_FFT_KERNELS=generic) and the two-level plane and
    z-stick kernels give the same H|psi> at cfg2 to rounding, in FP64 and in mixed precision."""
    case = kcase(W.cfg2(), 0)
    psi = W.random_psi(case.kp.kinetic_energy_squared, 8, seed=5)
    h_tl = _applicator(gpu_context, case)
    monkeypatch.setenv("Q3X_FFT_KERNELS", "generic")
    h_gen = _applicator(gpu_context, case)
    monkeypatch.delenv("Q3X_FFT_KERNELS")
    for prec, tol in ((FP64, 1e-13), (MIXED, 2e-6)):
        assert rel_err(h_tl.apply_batch(psi, prec), h_gen.apply_batch(psi, prec)) <= tol


@pytest.mark.parametrize("make,kindex", [(W.cfg1, 3), (W.cfg4, 1)])
def test_dense_grid_potential_is_restricted_to_the_wave_grid(gpu_context, make, kindex):
    """update_potential with the density-grid V on the reference's default wavefunction grid
    (2 x ecutwfc, smaller than the density grid): the library's G-space restriction on the device
    gives the H|psi> of the oracle's restriction (planewave_potential_utils.hpp:33-130), FP64
    <= 1e-10.  Feeding the dense V sampled at the coarse points instead is off by far more."""
    from q3_amd.planewave import PlanewaveBasis
    w = make()
    b = PlanewaveBasis(w.crystal, w.ecutwfc_ry, 0.0, False, w.fft_factors)
    assert b.rho_dims != b.wave_dims
    kp = b.compute_gk(w.kpoints[kindex])
    v_dense = W.local_potential(w.crystal, b, w.species, grid="rho")
    beta = kp.kb_projectors(w.crystal, w.species)
    bs, db = W.d_blocks_for(w.crystal, w.species)
    psi = W.random_psi(kp.kinetic_energy_squared, 6, seed=21)
    print(f"{w.name}: rho {b.rho_dims} wave {b.wave_dims} npw {kp.npw}", flush=True)
    h = PlanewaveHamiltonianApplicator(gpu_context, b)
    h.update_potential(v_dense)
    h.prepare_kpoint(kp)
    h.set_projectors(beta, bs, db)
    got = h.apply_batch(psi, FP64)
    v_s = O.restrict_potential(v_dense, b.wave_dims)
    ref = O.hamiltonian_apply(b.wave_dims, kp.fft_map, kp.kinetic_energy_squared, v_s, beta, bs, db, psi)
    assert rel_err(got, ref) <= 1e-10
    step = [r // s for r, s in zip(b.rho_dims, b.wave_dims)]
    if all(r == s * t for r, s, t in zip(b.rho_dims, b.wave_dims, step)):
        sampled = np.ascontiguousarray(v_dense[::step[0], ::step[1], ::step[2]])
        wrong = O.hamiltonian_apply(b.wave_dims, kp.fft_map, kp.kinetic_energy_squared, sampled, beta, bs, db, psi)
        assert rel_err(wrong, ref) > 1e-6
