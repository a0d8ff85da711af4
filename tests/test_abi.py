"""C-ABI checks that need no GPU: the library loads and exports every symbol include/q3x.h
declares; the host-side setup (G sphere, per-k lists, FFT maps, KB projectors) agrees with
the oracle bit for bit (maps) or to rounding (projectors); the host subspace eigensolver
agrees with scipy; argument errors come back as statuses with messages."""
import ctypes
import os
import re

import numpy as np
import pytest
import scipy.linalg

from oracle import oracle as O
from q3_amd import _lib
from q3_amd import workloads as W
from q3_amd.planewave import PlanewaveBasis

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared_symbols():
    text = open(os.path.join(ROOT, "include", "q3x.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(q3x_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    L = ctypes.CDLL(_lib.LIB_PATH)
    syms = _declared_symbols()
    assert len(syms) >= 25
    missing = [s for s in syms if not hasattr(L, s)]
    assert not missing, missing
    assert set(syms) == set(_lib.SIGNATURES), set(syms) ^ set(_lib.SIGNATURES)


def test_version_and_error_reporting():
    L = _lib.lib()
    assert b"gfx950" in L.q3x_version()
    h = ctypes.c_void_p()
    lat = np.eye(3).ravel()
    st = L.q3x_basis_create(lat.ctypes.data_as(ctypes.c_void_p), -1.0, 10.0, 0.0, 1, None, ctypes.byref(h))
    assert st == 1 and b"invalid" in L.q3x_last_error()
    with pytest.raises(_lib.Q3XError):
        _lib.check(st)


@pytest.mark.parametrize("make", [W.cfg1, W.cfg2, W.cfg4])
def test_setup_bit_exact_against_oracle(make):
    w = make()
    c = w.crystal
    b = PlanewaveBasis(c, w.ecutwfc_ry, 0.0, True, w.fft_factors)
    ob = O.Basis(c.lattice, c.alat, w.ecutwfc_ry, 0.0, True, w.fft_factors)
    assert b.wave_dims == ob.wave_dims and b.rho_dims == ob.rho_dims
    m, g, g2 = b.g_vectors()
    om, og, og2 = ob.gvectors()
    np.testing.assert_array_equal(m, om)
    np.testing.assert_array_equal(g, og)
    np.testing.assert_array_equal(g2, og2)
    ks = w.kpoints if len(w.kpoints) <= 64 else w.kpoints[:64]
    for k in ks[:: max(1, len(ks) // 8)]:
        kp = b.compute_gk(k)
        gi, ek, fm = ob.kpoint(k)
        np.testing.assert_array_equal(kp.g_vector_indices, gi)
        np.testing.assert_array_equal(kp.kinetic_energy_squared, ek)
        np.testing.assert_array_equal(kp.fft_map, fm)


@pytest.mark.parametrize("make,kidx", [(W.cfg1, 5), (W.cfg4, 2)])
def test_kb_projectors_against_oracle(make, kidx):
    w = make()
    c = w.crystal
    b = PlanewaveBasis(c, w.ecutwfc_ry, 0.0, True, w.fft_factors)
    ob = O.Basis(c.lattice, c.alat, w.ecutwfc_ry, 0.0, True, w.fft_factors)
    kp = b.compute_gk(w.kpoints[kidx])
    beta = kp.kb_projectors(c, w.species)
    obeta = ob.kb_projectors(w.kpoints[kidx], kp.g_vector_indices, c.positions, c.species, w.species)
    assert beta.shape == obeta.shape == (kp.npw, 4 * c.number_of_atoms)
    assert np.abs(beta - obeta).max() <= 1e-13 * np.abs(obeta).max()


def test_d_blocks_against_oracle():
    """Library D blocks (q3x_kb_d_blocks) equal the oracle's on a two-species cell whose species
    carry two betas in one l channel with an off-diagonal dion, and on the serving species."""
    rng = np.random.default_rng(11)
    a = rng.uniform(1, 3, (4, 4))
    sp0 = {"l": [0, 1, 1, 2], "dion_ry": a + a.T}
    sp1 = {"l": [1, 0], "dion_ry": np.array([[1.5, 0.3], [0.3, -2.0]])}
    c = W.Crystal(W.FCC.copy(), 10.0, rng.uniform(size=(5, 3)), np.array([0, 1, 1, 0, 1], np.int32))
    bs, db = W.d_blocks_for(c, [sp0, sp1])
    obs, odb = O.d_blocks(c.species, [sp0, sp1])
    np.testing.assert_array_equal(bs, obs)
    np.testing.assert_array_equal(db, odb)
    assert list(bs) == [12, 4, 4, 12, 4]
    w = W.cfg2()
    bs, db = W.d_blocks_for(w.crystal, w.species)
    obs, odb = O.d_blocks(w.crystal.species, w.species)
    np.testing.assert_array_equal(bs, obs)
    np.testing.assert_array_equal(db, odb)


def _hegv(H, S, nvec):
    n = H.shape[0]
    H = np.asfortranarray(H, np.complex128)
    S = np.asfortranarray(S, np.complex128)
    w = np.zeros(n)
    X = np.zeros((n, nvec), np.complex128, order="F")
    p = lambda a: a.ctypes.data_as(ctypes.c_void_p)  # noqa: E731
    st = _lib.lib().q3x_hegv(n, p(H), p(S), p(w), p(X), nvec)
    return st, w, X


@pytest.mark.parametrize("n,nvec", [(1, 1), (7, 3), (64, 32), (200, 100)])
def test_host_subspace_eigensolver_against_scipy(n, nvec):
    rng = np.random.default_rng(n)
    A = rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n))
    H = A + A.conj().T
    B = rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n))
    S = np.eye(n) + 0.1 * (B @ B.conj().T) / n
    st, w, X = _hegv(np.triu(H), np.triu(S), nvec)   # only the upper triangles are read
    assert st == 0
    np.testing.assert_allclose(w, scipy.linalg.eigh(H, S, eigvals_only=True), rtol=1e-11, atol=1e-11)
    np.testing.assert_allclose(X.conj().T @ S @ X, np.eye(nvec), atol=1e-11)
    np.testing.assert_allclose(H @ X, S @ X * w[:nvec], atol=1e-10 * max(1.0, np.abs(w).max()))


@pytest.mark.parametrize("n,nvec", [(48, 24), (320, 160)])
def test_host_subspace_eigensolver_degenerate_spectrum(n, nvec):
    """Exactly degenerate eigenvalues (symmetric crystals): eigenvectors stay orthonormal and
    span the eigenspaces (inverse iteration with in-cluster re-orthogonalisation)."""
    rng = np.random.default_rng(7)
    vals = np.repeat(np.linspace(-3.0, 40.0, n // 4), 4)[:n]      # every eigenvalue 4-fold
    Q, _ = np.linalg.qr(rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n)))
    H = (Q * vals) @ Q.conj().T
    H = 0.5 * (H + H.conj().T)
    st, w, X = _hegv(np.triu(H), np.eye(n, dtype=complex), nvec)
    assert st == 0
    np.testing.assert_allclose(w, np.sort(vals), atol=1e-11 * 40)
    np.testing.assert_allclose(X.conj().T @ X, np.eye(nvec), atol=1e-10)
    assert np.abs(H @ X - X * w[:nvec]).max() <= 1e-10 * 40


def test_host_subspace_eigensolver_rejects_indefinite_overlap():
    st, _, _ = _hegv(np.eye(3, dtype=complex), np.diag([1.0, -1.0, 1.0]).astype(complex), 1)
    assert st == 4


def test_random_starting_block_is_seeded_and_damped():
    """q3x_kpoint_random_psi: the same seed gives the same block, another seed another one, and
    the entries times (1 + |k+G|^2) have real and imaginary parts uniform in [-1, 1) (mean 0,
    E|z|^2 = 2/3, all within the unit square)."""
    w = W.cfg4()
    b = PlanewaveBasis(w.crystal, w.ecutwfc_ry, 0.0, True, w.fft_factors)
    kp = b.compute_gk(w.kpoints[3])
    a1, a2, c = kp.random_psi(6, 11), kp.random_psi(6, 11), kp.random_psi(6, 12)
    assert a1.shape == (kp.npw, 6) and a1.flags.f_contiguous
    assert np.array_equal(a1, a2) and not np.allclose(a1, c)
    z = a1 * (1.0 + kp.kinetic_energy_squared)[:, None]
    assert abs(z.mean()) < 0.02 and abs((np.abs(z) ** 2).mean() - 2.0 / 3.0) < 0.02
    assert np.abs(z.real).max() <= 1.0 and np.abs(z.imag).max() <= 1.0
