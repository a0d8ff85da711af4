"""The oracle and the library's host setup against the committed fixtures in tests/golden/:
the known answers of the reference's own unit tests (reference_known_answers.json, each entry
naming its source) and the dense cfg1 solution (cfg1_dense.npz, made by make_cfg1_dense.py)."""
import json
import os

import numpy as np
import pytest

from oracle import oracle as O
from q3_amd import workloads as W
from tests.helpers import check_eigenvalues

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
KA = json.load(open(os.path.join(GOLDEN, "reference_known_answers.json")))
HA = 2.0


def _lm(l, m):
    # combined index of harmonics.hpp:50-58: m = 0, +1, -1, +2, -2, ...
    return l * l if m == 0 else (l * l + 2 * m - 1 if m > 0 else l * l - 2 * m)


def test_good_fft_order_known_answers():
    for n, f, want in KA["good_fft_order"]["cases"]:
        assert O.good_fft_order(n, f) == want


def test_spherical_bessel_high_precision_table():
    e = KA["spherical_bessel_high_precision"]
    for l, x, want in e["cases"]:
        assert abs(O.sph_bessel(l, x) - want) <= e["relative_tolerance"] * abs(want), (l, x)


def test_spherical_bessel_batch_agrees_with_scalar():
    e = KA["spherical_bessel_batch_vs_scalar"]
    r = 1e-5 * np.exp(0.04 * np.arange(300))
    for l in e["degrees"]:
        for q in e["q_values"]:
            got = O.sph_bessel_batch(l, q, r)
            want = np.array([O.sph_bessel(l, q * x) for x in r])
            assert np.all(np.abs(got - want) <= e["relative_tolerance"] * np.abs(want)), (l, q)


def test_lagrange_known_answers():
    for table, q, dq, want, tol in KA["lagrange_4point"]["cases"]:
        got = O.lagrange4(table, q, dq)
        if tol == 0.0:
            assert got == want, (table, q)
        else:
            assert abs(got - want) <= tol, (table, q, got)


def test_simpson_known_answer():
    e = KA["simpson_r_squared"]
    r = 0.1 * np.arange(101)
    assert abs(O.simpson(r * r, np.full(101, 0.1)) - e["expected"]) <= e["tolerance"]


def test_ylm_known_answers_and_zero_vector_convention():
    e = KA["ylm_on_z_axis"]
    y = O.ylm(1, np.array([e["g"]]))[0]
    for l, m, want in e["cases"]:
        assert abs(y[_lm(l, m)] - want) <= e["tolerance"]
    e = KA["ylm_zero_vector_convention"]
    y = O.ylm(e["lmax"], np.zeros((1, 3)))[0]
    for l, m, want in e["exact"]:
        assert abs(y[_lm(l, m)] - want) <= e["tolerance"]
    for l, m, bound in e["negative"]:
        assert y[_lm(l, m)] < bound


def test_subspace_two_by_two_known_answer():
    e = KA["subspace_two_by_two"]
    info, w, X = O.zhegv(np.array(e["H"], complex), np.array(e["S"], complex))
    assert info == 0
    assert np.abs(w - np.array(e["eigenvalues"])).max() <= e["tolerance"]


# ------------------------------------------------------------------ cfg1 dense fixture ----
@pytest.fixture(scope="module")
def fx():
    return np.load(os.path.join(GOLDEN, "cfg1_dense.npz"))


@pytest.fixture(scope="module")
def oracle_cfg1():
    work = W.cfg1()
    c = work.crystal
    b = O.Basis(c.lattice, c.alat, work.ecutwfc_ry, 0.0, True, work.fft_factors)
    return work, b


def test_oracle_maps_match_fixture(fx, oracle_cfg1):
    work, b = oracle_cfg1
    assert tuple(fx["wave_dims"]) == tuple(b.wave_dims)
    for j, k in enumerate(fx["kpoints"]):
        gi, ek, fm = b.kpoint(k)
        np.testing.assert_array_equal(gi, fx[f"k{j}_g_index"])
        np.testing.assert_array_equal(ek, fx[f"k{j}_kinetic_ry"])
        np.testing.assert_array_equal(fm, fx[f"k{j}_fft_map"])


def test_library_setup_matches_fixture(fx):
    from q3_amd.planewave import PlanewaveBasis
    work = W.cfg1()
    b = PlanewaveBasis(work.crystal, work.ecutwfc_ry, 0.0, True, work.fft_factors)
    assert tuple(fx["wave_dims"]) == tuple(b.wave_dims)
    for j, k in enumerate(fx["kpoints"]):
        kp = b.compute_gk(k)
        np.testing.assert_array_equal(kp.g_vector_indices, fx[f"k{j}_g_index"])
        np.testing.assert_array_equal(kp.kinetic_energy_squared, fx[f"k{j}_kinetic_ry"])
        np.testing.assert_array_equal(kp.fft_map, fx[f"k{j}_fft_map"])


def _oracle_case(work, b, k):
    c = work.crystal
    gi, ek, fm = b.kpoint(k)
    m, g, g2 = b.gvectors()

    class _View:
        def g_vectors(self):
            return m, g, g2
        wave_dims = b.wave_dims
        cell_volume = abs(np.linalg.det(c.lattice)) * c.alat ** 3

    V = W.local_potential(c, _View(), work.species)
    beta = b.kb_projectors(k, gi, c.positions, c.species, work.species)
    bs, db = W.d_blocks_for(c, work.species)
    return fm, ek, V, beta, bs, db


def test_oracle_hamiltonian_matches_dense_fixture(fx, oracle_cfg1):
    work, b = oracle_cfg1
    fm, ek, V, beta, bs, db = _oracle_case(work, b, fx["kpoints"][0])
    got = O.hamiltonian_apply(b.wave_dims, fm, ek, V, beta, bs, db, np.asfortranarray(fx["k0_psi"]))
    want = fx["k0_hpsi"]
    assert np.linalg.norm(got - want) <= 1e-12 * np.linalg.norm(want)


@pytest.mark.parametrize("j", [0, 1, 2])
def test_oracle_davidson_matches_dense_fixture(fx, oracle_cfg1, j):
    work, b = oracle_cfg1
    fm, ek, V, beta, bs, db = _oracle_case(work, b, fx["kpoints"][j])
    eig, _, rep = O.davidson(b.wave_dims, fm, ek, V, beta, bs, db, work.nbands, work.noccupied, None,
                             tol=1e-10, max_iter=200)
    assert rep["unconverged"] == 0
    check_eigenvalues(eig, fx[f"k{j}_eigenvalues_ry"], work.noccupied, work.nbands, j)
