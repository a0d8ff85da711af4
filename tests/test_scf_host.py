"""Host side of the SCF loop (q3_amd.scf) and its oracle (oracle/scf_oracle.py), on the CPU.

Oracle pins (known answers and mathematics):
* Ewald energy of FCC silicon, a = 10.263048 bohr, Z = 4: -16.79593991 Ry, the reference's own
  known answer (tests/unit/.../electrostatic/test_ewald_summation.cpp:15-30), independent of the
  splitting parameter eta.
* PZ LDA: Slater exchange -(3/4)(9/(4 pi^2))^(1/3)/rs Ha; correlation from PZ81 Table XII at rs = 1
  and rs = 4 by hand; the reference test's total at rs = 1, -1.036 +- 0.01 Ry
  (test_libxc.cpp:19-37); v_xc = d(rho e_xc)/d rho by central differences on both branches.
* Fermi-Dirac: the electron count is met, occupations lie in [0, 2], -TS <= 0 and vanishes as
  sigma -> 0; fixed occupations fill nelec / 2 bands.
* Broyden: after a step the corrected residual is orthogonal, in the metric, to every stored
  residual difference (the secant property of Johnson's update); on a linear contraction the
  iteration converges.
* Hartree: E_H of a superposition of narrow Gaussian charges plus the Ewald energy identity.

Product against oracle on the same inputs: LDA, Hartree, ionic potential, Ewald, occupations and
a Broyden sequence to 1e-12.
"""
import math

import numpy as np
import pytest

from oracle import scf_oracle as S
from q3_amd import scf as P
from q3_amd import workloads as W

FCC = np.array([[-0.5, 0.0, 0.5], [0.0, 0.5, 0.5], [-0.5, 0.5, 0.0]])


@pytest.mark.parametrize("eta", [None, 0.25, 0.5, 0.9])
def test_oracle_ewald_silicon_known_answer(eta):
    e = S.ewald(FCC, 10.263048, [[0, 0, 0], [0.25, 0.25, 0.25]], [4, 4], eta=eta)
    assert abs(e - (-16.79593991)) < 1e-6


def test_oracle_pz_known_values():
    rs = np.array([0.5, 1.0, 4.0])
    rho = 3.0 / (4.0 * math.pi * rs ** 3)
    exc, vxc = S.lda_pz(rho)
    ex_ha = -0.75 * (9.0 / (4.0 * math.pi ** 2)) ** (1.0 / 3.0) / rs
    ec_1 = -0.1423 / (1 + 1.0529 + 0.3334)
    ec_4 = -0.1423 / (1 + 1.0529 * 2.0 + 0.3334 * 4.0)
    ec_05 = 0.0311 * math.log(0.5) - 0.048 + 0.0020 * 0.5 * math.log(0.5) - 0.0116 * 0.5
    assert np.allclose(exc, 2.0 * (ex_ha + np.array([ec_05, ec_1, ec_4])), rtol=1e-13, atol=0)
    assert abs(exc[1] - (-1.036)) < 0.01
    # v = d(rho e)/d rho on both branches
    for r in (0.3, 0.9, 1.7, 6.0):
        rho0 = 3.0 / (4.0 * math.pi * r ** 3)
        h = rho0 * 1e-5
        ep, _ = S.lda_pz(np.array([rho0 + h]))
        em, _ = S.lda_pz(np.array([rho0 - h]))
        _, v = S.lda_pz(np.array([rho0]))
        num = ((rho0 + h) * ep[0] - (rho0 - h) * em[0]) / (2 * h)
        assert abs(num - v[0]) < 1e-8 * abs(v[0])


def test_oracle_fermi_dirac_properties():
    rng = np.random.default_rng(0)
    eig = np.sort(rng.uniform(-1, 1, size=(5, 8)), axis=1)
    wk = np.full(5, 0.2)
    for sigma in (0.05, 0.01):
        occ, ef, ts = S.fermi_dirac_occupations(eig, wk, 8, sigma)
        assert abs(np.sum(wk[:, None] * occ) - 8) < 1e-10
        assert occ.min() >= 0 and occ.max() <= 2 and ts <= 0
    occ, ef, ts = S.fermi_dirac_occupations(eig, wk, 8, 1e-6)
    assert abs(ts) < 1e-4
    occ0, _, ts0 = S.fermi_dirac_occupations(eig, wk, 8, 0.0)
    assert ts0 == 0 and np.all(occ0[:, :4] == 2) and np.all(occ0[:, 4:] == 0)


def test_oracle_broyden_secant_property_and_convergence():
    """On a linear map out(x) = M x + b each secant pair satisfies dF = (I - M) drho, so the
    corrected residual F - sum gamma dF is the true residual at x - sum gamma drho; it is
    metric-orthogonal to every stored dF, and the step is that point plus alpha times it.  A sign
    slip in either difference breaks the first identity."""
    rng = np.random.default_rng(1)
    n = 40
    w = rng.uniform(0.5, 2.0, n)
    M = rng.standard_normal((n, n)) * 0.08
    b = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    mixer = S.Broyden(w)
    x = np.zeros(n, np.complex128)
    for it in range(40):
        out = M @ x + b
        x_next = mixer.mix(x, out)
        F = out - x
        if mixer.dF:
            beta = np.array([[mixer.dot(p, q) for q in mixer.dF] for p in mixer.dF])
            gam = np.linalg.solve(beta, [mixer.dot(p, F) for p in mixer.dF])
            xc = x - sum(g * d for g, d in zip(gam, mixer.drho))
            Fc = F - sum(g * d for g, d in zip(gam, mixer.dF))
            assert np.abs(Fc - (M @ xc + b - xc)).max() < 1e-10 * max(1.0, np.abs(F).max())
            assert max(abs(mixer.dot(p, Fc)) for p in mixer.dF) < 1e-9 * max(1.0, mixer.dot(F, F))
            assert np.abs(x_next - (xc + 0.7 * Fc)).max() < 1e-12 * max(1.0, np.abs(x).max())
        x = x_next
    assert np.linalg.norm(M @ x + b - x) < 1e-6


def test_oracle_hartree_of_gaussians_against_ewald():
    """Narrow Gaussian charges Z (width s) at the ion sites in a neutralising background: E_H
    (G != 0) equals the Ewald energy of the point charges plus e^2 eta / sqrt(pi) sum Z^2 and
    e^2 pi Z_tot^2 / (2 Omega eta^2), eta = 1 / (2 s)."""
    lat, alat = FCC, 10.26
    pos = np.array([[0, 0, 0], [0.25, 0.25, 0.25]])
    dims = (48, 48, 48)
    s = 0.45
    G, m, omega = S.box_gvectors(lat, alat, dims)
    g2 = np.sum(G * G, axis=-1)
    rg = np.zeros(dims, np.complex128)
    for x in pos:
        rg += 4.0 / omega * np.exp(-g2 * s * s / 2) * np.exp(-2j * math.pi * (m @ x))
    rho = (np.fft.ifftn(rg) * rg.size).real
    _v, eh = S.hartree(rho, lat, alat)
    ew = S.ewald(lat, alat, pos, [4, 4])
    # the reciprocal Ewald term with eta = 1 / (2 s) is E_H of these Gaussians; what remains is
    # the self and background terms (the real-space term is erfc(eta d) ~ 1e-11 here)
    eta = 1.0 / (2.0 * s)
    self_e = 2.0 * eta / math.sqrt(math.pi) * 32.0
    bg = 2.0 * math.pi * 64.0 / (2.0 * omega * eta * eta)
    assert abs(eh - ew - self_e - bg) < 1e-8


def test_product_host_pieces_equal_the_oracle():
    w = W.cfg1()
    c = w.crystal
    dims = (30, 30, 30)
    rec = P.BoxReciprocal(c, dims)
    rng = np.random.default_rng(2)
    rho = rng.uniform(0.001, 0.2, size=dims)
    e1, v1 = P.pz_lda(rho)
    e2, v2 = S.lda_pz(rho)
    assert np.allclose(e1, e2, rtol=1e-13, atol=0) and np.allclose(v1, v2, rtol=1e-13, atol=0)
    vh_o, eh_o = S.hartree(rho, c.lattice, c.alat)
    rg = rec.to_g(rho)
    vh_p = rec.to_r(4 * math.pi * 2.0 * rg * rec.inv_g2)
    assert np.abs(vh_p - vh_o).max() < 1e-12 * np.abs(vh_o).max()
    sp = w.species
    vl_p = P.ionic_local_potential(c, sp, rec)
    vl_o = S.ionic_local(c.lattice, c.alat, dims, c.positions, [4.0, 4.0], [sp[0]["r_loc"]] * 2)
    assert np.abs(vl_p - vl_o).max() < 1e-12 * np.abs(vl_o).max()
    assert abs(P.ewald_energy(c, [4.0, 4.0]) - S.ewald(c.lattice, c.alat, c.positions, [4.0, 4.0])) < 1e-10
    eig = np.sort(rng.uniform(-1, 1, size=(6, 8)), axis=1)
    wk = np.full(6, 1 / 6)
    a, b = P.fermi_dirac(eig, wk, 8, 0.01), S.fermi_dirac_occupations(eig, wk, 8, 0.01)
    assert np.allclose(a[0], b[0], atol=1e-12) and abs(a[1] - b[1]) < 1e-12 and abs(a[2] - b[2]) < 1e-12
    # a Broyden sequence
    n = rho.size
    wts = rng.uniform(0.1, 1, n)
    pm, om = P.BroydenMixer(wts), S.Broyden(wts)
    x = np.zeros(n, np.complex128)
    M = rng.standard_normal(n) * 0.3
    for _ in range(12):
        out = M * x + 1.0
        xp = pm.mix(x, out)
        xo = om.mix(x, out)
        assert np.abs(xp - xo).max() < 1e-12
        x = xp


def test_oracle_density_two_plane_waves_closed_form():
    """psi = a e^{i(k+G1).r} + b e^{i(k+G2).r}: |psi(r)|^2 = |a|^2 + |b|^2 + 2 Re(a b* e^{i(G1-G2).r})
    at the box points r = (i1/n1, i2/n2, i3/n3) in crystal coordinates, which fixes the box's axis
    order and the sign of the transform."""
    from oracle import oracle as O
    w = W.cfg1()
    ob = O.Basis(w.crystal.lattice, w.crystal.alat, 8.0, 0.0, True, (1, 1, 1))
    k = w.kpoints[3]
    gi, ek, fm = ob.kpoint(k)
    m = ob.gvectors()[0][gi]
    dims = ob.wave_dims
    i, j = 3, 17
    a, b = 0.6 + 0.2j, -0.3 + 0.5j
    psi = np.zeros((len(gi), 1), np.complex128)
    psi[i, 0], psi[j, 0] = a, b
    rho = S.density(dims, fm, psi, [0.7])
    x = np.stack(np.meshgrid(*[np.arange(n) / n for n in dims], indexing="ij"), -1)
    ph = np.exp(2j * math.pi * (x @ (m[i] - m[j]).astype(float)))
    want = 0.7 * (abs(a) ** 2 + abs(b) ** 2 + 2 * (a * np.conj(b) * ph).real)
    assert np.abs(rho - want).max() < 1e-13
    # Parseval: the box mean of |psi|^2 is sum |c|^2
    rng = np.random.default_rng(3)
    p2 = rng.standard_normal((len(gi), 3)) + 1j * rng.standard_normal((len(gi), 3))
    r2 = S.density(dims, fm, p2, [1.0, 2.0, 0.5])
    assert abs(r2.mean() - float(np.sum(np.abs(p2) ** 2 * np.array([1.0, 2.0, 0.5])))) < 1e-9 * r2.mean()
