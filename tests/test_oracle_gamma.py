"""The oracle's Gamma trick (half sphere, two real bands per complex transform, 2 Re(.) less the
G = 0 term in every plane-wave product, real subspace pencil) pinned against its own full-sphere
path, which the dense brute-force and closed-form tests pin.

A half-sphere vector psi stands for the real function whose full-sphere coefficients are psi(G)
on the half and conj(psi(G)) at -G; H maps real functions to real functions (V(r) real, each KB
projector real in real space: beta(-G) = conj(beta(G))), so H_gamma psi must equal the half-sphere
part of H_full psi_full, and the Gamma Davidson must find the full-sphere spectrum.
"""
import math

import numpy as np
import pytest

from oracle import oracle as O
from q3_amd import workloads as W


class _OB:
    """An oracle basis seen through the attributes workloads.local_potential reads."""

    def __init__(self, ob, crystal):
        self.ob = ob
        self.wave_dims = ob.wave_dims
        self.gamma_only = ob.gamma_only
        _r, self.cell_volume, _t = O.crystal(crystal.lattice, crystal.alat)

    def g_vectors(self):
        return self.ob.gvectors()


def _pair(ecut_ry=16.0, exact=True):
    w = W.cfg1()
    c = w.crystal
    full = O.Basis(c.lattice, c.alat, ecut_ry, 0.0, exact, w.fft_factors)
    half = O.Basis(c.lattice, c.alat, ecut_ry, 0.0, exact, w.fft_factors, gamma_only=True)
    return w, c, full, half


def test_half_sphere_is_the_reference_rule():
    """g_vectors_grid.hpp:69-76: i3 > 0, or i3 = 0 and i2 > 0, or i3 = i2 = 0 and i1 >= 0; with the
    mirror images it is the full sphere; sorted by |G|^2."""
    _, _, full, half = _pair()
    mf, _, g2f = full.gvectors()
    mh, _, g2h = half.gvectors()
    keep = (mf[:, 2] > 0) | ((mf[:, 2] == 0) & (mf[:, 1] > 0)) | ((mf[:, 2] == 0) & (mf[:, 1] == 0) & (mf[:, 0] >= 0))
    assert {tuple(x) for x in mh} == {tuple(x) for x in mf[keep]}
    assert len(mh) == (len(mf) + 1) // 2
    both = {tuple(x) for x in mh} | {tuple(-x) for x in mh}
    assert both == {tuple(x) for x in mf}
    assert np.all(np.diff(g2h) >= 0) and tuple(mh[0]) == (0, 0, 0)


def test_gamma_kpoint_list_and_mirror_map():
    _, _, full, half = _pair()
    gf, ekf, fmf = full.kpoint([0, 0, 0])
    gh, ekh, fmh = half.kpoint([0, 0, 0])
    mm = half.minus_map([0, 0, 0])
    assert len(gh) == (len(gf) + 1) // 2 and ekh[0] == 0.0
    mh, _, _ = half.gvectors()
    d = np.array(half.wave_dims)
    for j in (0, 1, 7, len(gh) - 1):
        m = -mh[gh[j]]
        w = np.where(m < 0, m + d, m)
        assert mm[j] == (w[0] * d[1] + w[1]) * d[2] + w[2]
    assert len(set(fmh) | set(mm)) == len(gf)  # +G and -G boxes cover the full list once (G = 0 shared)


def _real_function_cases(ecut_ry=16.0):
    w, c, full, half = _pair(ecut_ry)
    k = [0.0, 0.0, 0.0]
    gf, ekf, fmf = full.kpoint(k)
    gh, ekh, fmh = half.kpoint(k)
    mm = half.minus_map(k)
    mf, _, _ = full.gvectors()
    mh, _, _ = half.gvectors()
    pos = {tuple(mf[g]): i for i, g in enumerate(gf)}
    plus = np.array([pos[tuple(mh[g])] for g in gh])
    minus = np.array([pos[tuple(-mh[g])] for g in gh])
    V = W.local_potential(c, _OB(full, c), w.species)
    Vh = W.local_potential(c, _OB(half, c), w.species)
    beta_f = full.kb_projectors(k, gf, c.positions, c.species, w.species)
    beta_h = half.kb_projectors(k, gh, c.positions, c.species, w.species)
    bs, db = O.d_blocks(c.species, w.species)
    return dict(full=full, half=half, gf=gf, ekf=ekf, fmf=fmf, gh=gh, ekh=ekh, fmh=fmh, mm=mm, plus=plus,
                minus=minus, V=V, Vh=Vh, beta_f=beta_f, beta_h=beta_h, bs=bs, db=db)


def test_gamma_inputs_describe_real_functions():
    t = _real_function_cases()
    assert np.abs(t["V"] - t["Vh"]).max() <= 1e-12 * np.abs(t["V"]).max()
    bf, bh = t["beta_f"], t["beta_h"]
    assert np.abs(bf[t["plus"]] - bh).max() <= 1e-13 * np.abs(bf).max()
    assert np.abs(bf[t["minus"]] - np.conj(bh)).max() <= 1e-13 * np.abs(bf).max()


@pytest.mark.parametrize("nvec", [1, 4, 5])
def test_gamma_hpsi_equals_full_sphere_on_real_functions(nvec):
    """Pairs and the odd last band alone (serial.hpp:283-418), the projector term's 2 Re(.) less
    G = 0 (serial.hpp:519-566).  A dropped G = 0 correction, a swapped pair or a missing 1/2 in
    the separation each fail by O(1)."""
    t = _real_function_cases()
    rng = np.random.default_rng(5)
    nh = len(t["gh"])
    psi = (rng.standard_normal((nh, nvec)) + 1j * rng.standard_normal((nh, nvec))) / (1 + t["ekh"])[:, None]
    psi[0, :] = psi[0, :].real
    full = np.zeros((len(t["gf"]), nvec), np.complex128)
    full[t["minus"]] = np.conj(psi)
    full[t["plus"]] = psi
    hf = O.hamiltonian_apply(t["full"].wave_dims, t["fmf"], t["ekf"], t["V"], t["beta_f"], t["bs"], t["db"], full)
    hg = O.hamiltonian_apply(t["half"].wave_dims, t["fmh"], t["ekh"], t["V"], t["beta_h"], t["bs"], t["db"], psi,
                             minus_map=t["mm"])
    want = hf[t["plus"]]
    assert np.abs(hg - want).max() <= 1e-12 * np.abs(want).max()
    # the full-sphere image is itself real: H psi at -G is the conjugate
    assert np.abs(hf[t["minus"]] - np.conj(want)).max() <= 1e-12 * np.abs(want).max()


def test_gamma_davidson_finds_the_full_sphere_spectrum():
    """Both start from damped random blocks (plane-wave unit vectors seed a Gamma solve with cosines
    only and can miss a symmetry sector of this symmetric cell)."""
    t = _real_function_cases()
    nb = 8
    p0f = W.random_psi(t["ekf"], nb, seed=3)
    p0g = W.random_psi(t["ekh"], nb, seed=4)
    p0g[0, :] = p0g[0, :].real
    ef, _, rf = O.davidson(t["full"].wave_dims, t["fmf"], t["ekf"], t["V"], t["beta_f"], t["bs"], t["db"], nb, nb,
                           p0f, tol=1e-11, max_iter=200)
    eg, psig, rg = O.davidson(t["half"].wave_dims, t["fmh"], t["ekh"], t["V"], t["beta_h"], t["bs"], t["db"], nb, nb,
                              p0g, tol=1e-11, max_iter=200, minus_map=t["mm"])
    assert rf["unconverged"] == 0 and rg["unconverged"] == 0
    assert np.abs(eg - ef).max() <= 1e-9
    assert np.all(psig[0].imag == 0.0)  # G = 0 kept real
    # real-function orthonormality: 2 Re(psi^H psi) - psi0^* psi0 = I
    G = 2 * (psig.conj().T @ psig).real - np.outer(psig[0].conj(), psig[0]).real
    assert np.abs(G - np.eye(nb)).max() <= 1e-10
    # Rayleigh quotients of the returned vectors are the eigenvalues
    hp = O.hamiltonian_apply(t["half"].wave_dims, t["fmh"], t["ekh"], t["V"], t["beta_h"], t["bs"], t["db"], psig,
                             minus_map=t["mm"])
    rq = 2 * np.einsum("gi,gi->i", psig.conj(), hp).real - (psig[0].conj() * hp[0]).real
    assert np.abs(rq - eg).max() <= 1e-9
