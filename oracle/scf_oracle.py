"""CPU restatement of the self-consistent field loop around the H|psi> + Davidson path.

Test infrastructure only (tests/ load it as the checker of q3_amd.scf).  It follows the
reference's SCF iteration (src/layer_4_computation/iterator/init.hpp:82-145,
iterator/kohn_sham/init.hpp:357-470, iterator/kohn_sham/planewave/init.hpp:869-1091) on the
pieces this path needs, each with the reference or published source it restates:

  density       PlanewaveChargeDensityAccumulator (charge_density/planewave/internal/
                accumulator.hpp:77-280): rho(r) = sum_k w_k sum_n f_nk |psi_nk(r)|^2 / Omega, the
                band transformed with the unnormalised backward FFT
  Hartree       V_H(G) = 4 pi e^2 rho(G) / |G|^2, G != 0, e^2 = 2 (Rydberg units)
  LDA           Perdew-Zunger 1981 (PRB 23, 5048, Table XII) fit of Ceperley-Alder, Slater
                exchange; the functional "PZ" of the reference (test_libxc.cpp:19-37)
  Ewald         the Ewald sum of the point ions in a neutralising background
                (effective_potential/internal/electrostatic/ewald_summation.hpp)
  mixing        Johnson's modified Broyden in the Hartree metric, history 8, alpha 0.7
                (iterator/internal/mixing/reciprocal_space/broyden.hpp:64-380), no preconditioner
  occupations   Fermi-Dirac smearing f(x) = 1 / (e^x + 1), x = (e - E_F) / sigma, two electrons per
                band, E_F by bisection to the electron count, and the -TS term
                sigma sum w f ln f + (1 - f) ln(1 - f) (iterator/internal/smearing/
                fermi_dirac.hpp:19-69); sigma = 0 gives fixed occupations
  tolerance     Davidson threshold 1e-2 Ry on iteration 1, then min(tol, 0.1 dr2 / N_el), floor
                1e-13 (kohn_sham/planewave/init.hpp:869-907); the first-iteration
                re-diagonalisation when dr2 < tol N_el (kohn_sham/init.hpp:407-430)
  energy        the Harris-Foulkes form E = E_band - <rho_in|V_Hxc> + E_H + E_xc + E_Ewald on the
                input density, which equals the reference's total energy at self-consistency

The ions' local potential is a Gaussian-smeared point charge, V(G) = -4 pi Z e^2
exp(-|G|^2 r_loc^2 / 2) / (Omega |G|^2), with its finite G = 0 part 2 pi Z e^2 r_loc^2 / Omega.
"""
from __future__ import annotations

import math

import numpy as np
from scipy.special import erfc

from oracle import oracle as O

E2 = 2.0  # e^2 in Rydberg atomic units


# ---- density
def density(dims, fftmap, psi, weights, minus_map=None):
    """sum_b weights[b] |psi_b(r)|^2 on the box, C order.  psi_b(r) = sum_G c_b(G) e^{i(k+G).r}
    (numpy's inverse FFT times N); with a Gamma-trick minus map, c(-G) = conj(c(G))."""
    n = int(np.prod(dims))
    out = np.zeros(n)
    psi = np.asarray(psi).reshape(len(fftmap), -1)
    for b in range(psi.shape[1]):
        box = np.zeros(n, np.complex128)
        box[np.asarray(fftmap)] = psi[:, b]
        if minus_map is not None:
            box[np.asarray(minus_map)] = np.conj(psi[:, b])
            box[np.asarray(fftmap)] = psi[:, b]
        f = np.fft.ifftn(box.reshape(dims)) * n
        out += weights[b] * (f.real ** 2 + f.imag ** 2).ravel()
    return out.reshape(dims)


# ---- exchange-correlation
PZ = dict(gamma=-0.1423, beta1=1.0529, beta2=0.3334, A=0.0311, B=-0.048, C=0.0020, D=-0.0116)


def lda_pz(rho):
    """Energy per electron and potential (Ry) of the PZ81 LDA for densities rho (bohr^-3)."""
    rho = np.asarray(rho, dtype=np.float64)
    exc = np.zeros_like(rho)
    vxc = np.zeros_like(rho)
    m = rho > 1e-30
    rs = (3.0 / (4.0 * math.pi * rho[m])) ** (1.0 / 3.0)
    # Slater exchange in Hartree: e_x = -(3/4)(3/pi)^(1/3) rho^(1/3) = -0.458165.../rs
    cx = 0.75 * (3.0 / (2.0 * math.pi)) ** (2.0 / 3.0)
    ex = -cx / rs
    vx = 4.0 / 3.0 * ex
    ec = np.empty_like(rs)
    vc = np.empty_like(rs)
    hi = rs >= 1.0
    s = np.sqrt(rs[hi])
    den = 1.0 + PZ["beta1"] * s + PZ["beta2"] * rs[hi]
    ec[hi] = PZ["gamma"] / den
    vc[hi] = ec[hi] * (1.0 + 7.0 / 6.0 * PZ["beta1"] * s + 4.0 / 3.0 * PZ["beta2"] * rs[hi]) / den
    lo = ~hi
    r, lr = rs[lo], np.log(rs[lo])
    ec[lo] = PZ["A"] * lr + PZ["B"] + PZ["C"] * r * lr + PZ["D"] * r
    vc[lo] = PZ["A"] * lr + (PZ["B"] - PZ["A"] / 3.0) + 2.0 / 3.0 * PZ["C"] * r * lr + \
        (2.0 * PZ["D"] - PZ["C"]) / 3.0 * r
    exc[m] = 2.0 * (ex + ec)  # Hartree -> Rydberg
    vxc[m] = 2.0 * (vx + vc)
    return exc, vxc


# ---- reciprocal-space helpers on the full box
def box_gvectors(lattice_rows, alat, dims):
    """Cartesian G (bohr^-1) of every box point, C order, Miller indices wrapped to the centre."""
    A = np.asarray(lattice_rows, float).reshape(3, 3) * alat
    B = 2 * math.pi * np.linalg.inv(A).T  # rows b_i
    ms = [np.fft.fftfreq(n, 1.0 / n) for n in dims]
    m1, m2, m3 = np.meshgrid(*ms, indexing="ij")
    G = m1[..., None] * B[0] + m2[..., None] * B[1] + m3[..., None] * B[2]
    omega = abs(np.linalg.det(A))
    return G, np.stack([m1, m2, m3], axis=-1), omega


def hartree(rho_r, lattice_rows, alat):
    """V_H(r) (Ry) and E_H (Ry) of a periodic density, G = 0 left out."""
    dims = rho_r.shape
    G, _m, omega = box_gvectors(lattice_rows, alat, dims)
    g2 = np.sum(G * G, axis=-1)
    rg = np.fft.fftn(rho_r) / rho_r.size
    vg = np.zeros_like(rg)
    nz = g2 > 1e-12
    vg[nz] = 4 * math.pi * E2 * rg[nz] / g2[nz]
    eh = 0.5 * omega * float(np.sum((np.conj(rg) * vg).real))
    vr = np.fft.ifftn(vg).real * rho_r.size
    return vr, eh


def ionic_local(lattice_rows, alat, dims, positions, charges, r_locs):
    """Local potential (Ry) of Gaussian-smeared ions on the box, with the finite G = 0 part."""
    G, m, omega = box_gvectors(lattice_rows, alat, dims)
    g2 = np.sum(G * G, axis=-1)
    vg = np.zeros(dims, np.complex128)
    nz = g2 > 1e-12
    pos = np.asarray(positions, float)
    g0 = 0.0
    for x, z, rl in zip(pos, charges, r_locs):
        phase = np.exp(-2j * math.pi * (m @ x))
        form = np.zeros(dims)
        form[nz] = -4 * math.pi * z * E2 * np.exp(-g2[nz] * rl * rl / 2) / (omega * g2[nz])
        vg += form * phase
        g0 += 2 * math.pi * z * E2 * rl * rl / omega
    vg[0, 0, 0] = g0
    return np.fft.ifftn(vg).real * vg.size


def ewald(lattice_rows, alat, positions, charges, eta=None, tol=1e-14):
    """Ewald energy (Ry) of point charges in a neutralising background."""
    A = np.asarray(lattice_rows, float).reshape(3, 3) * alat
    omega = abs(np.linalg.det(A))
    B = 2 * math.pi * np.linalg.inv(A).T
    tau = np.asarray(positions, float) @ A
    z = np.asarray(charges, float)
    ztot = z.sum()
    if eta is None:
        eta = math.sqrt(math.pi) / omega ** (1.0 / 3.0)
    # real space: sum over pairs and images of z_i z_j erfc(eta r) / r
    rcut = math.sqrt(-math.log(tol)) / eta + 1.0
    nmax = [int(math.ceil(rcut / (omega / np.linalg.norm(np.cross(A[(i + 1) % 3], A[(i + 2) % 3]))))) + 1
            for i in range(3)]
    er = 0.0
    rng = [np.arange(-n, n + 1) for n in nmax]
    cells = np.array(np.meshgrid(*rng, indexing="ij")).reshape(3, -1).T @ A
    for i in range(len(z)):
        for j in range(len(z)):
            d = tau[i] - tau[j] + cells
            r = np.linalg.norm(d, axis=1)
            keep = r > 1e-10
            er += z[i] * z[j] * float(np.sum(erfc(eta * r[keep]) / r[keep]))
    er *= 0.5
    # reciprocal space
    gmax = 2 * eta * math.sqrt(-math.log(tol))
    mmax = [int(math.ceil(gmax / np.linalg.norm(B[i]))) + 1 for i in range(3)]
    rngg = [np.arange(-n, n + 1) for n in mmax]
    mm = np.array(np.meshgrid(*rngg, indexing="ij")).reshape(3, -1).T
    mm = mm[np.any(mm != 0, axis=1)]
    Gv = mm @ B
    g2 = np.sum(Gv * Gv, axis=1)
    s = np.exp(1j * Gv @ tau.T) @ z
    eg = 2 * math.pi / omega * float(np.sum(np.exp(-g2 / (4 * eta * eta)) / g2 * np.abs(s) ** 2))
    eself = -eta / math.sqrt(math.pi) * float(np.sum(z * z))
    ebg = -math.pi * ztot * ztot / (2 * omega * eta * eta)
    return E2 * (er + eg + eself + ebg)  # e^2 = 2 in Ry


# ---- mixing
class Broyden:
    """Johnson's modified Broyden in the Hartree metric (broyden.hpp:64-380): w(G) the metric
    weights (G = 0 excluded by a zero weight), alpha the linear part, history depth 8."""

    def __init__(self, weights, alpha=0.7, depth=8):
        self.w = np.asarray(weights, float).ravel()
        self.alpha = alpha
        self.depth = depth
        self.it = 0
        self.dF, self.drho = [], []
        self.prev = None

    def dot(self, a, b):
        return float(np.sum(self.w * (np.conj(a) * b).real))

    def mix(self, rin, rout):
        """rin, rout: flat complex rho(G); returns the next input."""
        self.it += 1
        rin = rin.copy()
        F = rout - rin
        if self.it > 1 and self.prev is not None:
            Fp, rp = self.prev
            dF, dr = Fp - F, rp - rin
            if len(self.dF) < self.depth:
                self.dF.append(dF)
                self.drho.append(dr)
            else:
                pos = (self.it - 2) % self.depth
                self.dF[pos], self.drho[pos] = dF, dr
        self.prev = (F.copy(), rin.copy())
        if min(self.it - 1, self.depth) > 0 and self.dF:
            n = len(self.dF)
            beta = np.array([[self.dot(self.dF[i], self.dF[j]) for j in range(n)] for i in range(n)])
            proj = np.array([self.dot(self.dF[i], F) for i in range(n)])
            gam = np.linalg.solve(beta, proj)
            for i in range(n):
                rin -= gam[i] * self.drho[i]
                F -= gam[i] * self.dF[i]
        return rin + self.alpha * F


# ---- occupations
def fermi_dirac_occupations(eig, kweights, nelec, sigma):
    """(occupations nk x nb in electrons per band, Fermi level, -TS in Ry)."""
    eig = np.asarray(eig, float)
    wk = np.asarray(kweights, float)[:, None]
    if sigma <= 0.0:
        occ = np.zeros_like(eig)
        occ[:, :nelec // 2] = 2.0
        return occ, float(np.max(eig[:, :nelec // 2])), 0.0

    def count(ef):
        with np.errstate(over="ignore"):
            return float(np.sum(wk * 2.0 / (np.exp((eig - ef) / sigma) + 1.0)))
    lo, hi = float(eig.min()) - 10 * sigma, float(eig.max()) + 10 * sigma
    for _ in range(200):
        mid = 0.5 * (lo + hi)
        if count(mid) < nelec:
            lo = mid
        else:
            hi = mid
        if hi - lo < 1e-15:
            break
    ef = 0.5 * (lo + hi)
    with np.errstate(over="ignore"):
        f = 1.0 / (np.exp((eig - ef) / sigma) + 1.0)
    with np.errstate(divide="ignore", invalid="ignore"):
        t1 = np.where(f > 1e-16, f * np.log(f), 0.0)
        t0 = np.where(1 - f > 1e-16, (1 - f) * np.log(1 - f), 0.0)
    ts = float(np.sum(wk * 2.0 * sigma * (t1 + t0)))
    return 2.0 * f, ef, ts


# ---- the loop
def scf(lattice_rows, alat, positions, species_index, species, kpoints, kweights, ecutwfc, nbands, nelec,
        fft_factors=(1, 1, 1), conv_thr=1e-10, max_iter=60, davidson_iter=40, sigma=0.01, log=None):
    """Self-consistent cycle, Fermi-Dirac occupations of width sigma (Ry).  Returns a dict
    with the total energy (Ry), its parts, the eigenvalues (nk x nb, Ry), the density and the
    iteration record."""
    ob = O.Basis(lattice_rows, alat, ecutwfc, 0.0, True, fft_factors)
    dims = tuple(ob.wave_dims)
    assert tuple(ob.rho_dims) == dims, "the loop runs on one box (wavefunction box = density box)"
    G, _m, omega = box_gvectors(lattice_rows, alat, dims)
    g2 = np.sum(G * G, axis=-1).ravel()
    w = np.zeros_like(g2)
    nz = g2 > 1e-12
    w[nz] = E2 * 4 * math.pi * omega * 0.5 / g2[nz]  # Hartree metric (kohn_sham/planewave/init.hpp:940-946)
    zs = [species[s]["z_val"] for s in species_index]
    rls = [species[s]["r_loc"] for s in species_index]
    vloc = ionic_local(lattice_rows, alat, dims, positions, zs, rls)
    e_ewald = ewald(lattice_rows, alat, positions, zs)
    bs, db = O.d_blocks(species_index, species)
    kd = []
    for k in kpoints:
        gi, ek, fm = ob.kpoint(k)
        beta = ob.kb_projectors(k, gi, positions, species_index, species)
        kd.append((ek, fm, beta))
    rho_in = np.full(dims, nelec / omega)
    psis = [None] * len(kpoints)
    mixer = Broyden(w)
    tol = 1e-2
    hist = []
    eig = np.zeros((len(kpoints), nbands))
    rediag_done = False
    for it in range(1, max_iter + 1):
        vh, eh = hartree(rho_in, lattice_rows, alat)
        exc, vxc = lda_pz(rho_in)
        exc_tot = omega / rho_in.size * float(np.sum(exc * rho_in))
        vhxc = vh + vxc
        veff = vloc + vhxc
        if it > 1:
            tol = max(min(tol, 0.1 * hist[-1]["dr2"] / max(1.0, nelec)), 1e-13)

        def solve_all(tol_):
            for i, (ek, fm, beta) in enumerate(kd):
                psi0 = psis[i] if psis[i] is not None else O_random(ek, nbands, i)
                # every band held to the threshold: under smearing the bands above nelec / 2 carry
                # charge too
                e, p, _rep = O.davidson(dims, fm, ek, veff.ravel(), beta, bs, db, nbands, nbands, psi0=psi0,
                                        tol=tol_, max_iter=davidson_iter)
                eig[i], psis[i] = e, p
            occ, ef, ts = fermi_dirac_occupations(eig, kweights, nelec, sigma)
            rho = np.zeros(rho_in.size)
            for i, (ek, fm, beta) in enumerate(kd):
                rho += density(dims, fm, psis[i], kweights[i] * occ[i] / omega).ravel()
            return rho.reshape(dims), occ, ef, ts

        rho_out, occ, ef, ts = solve_all(tol)
        rg_in = np.fft.fftn(rho_in).ravel() / rho_in.size
        rg_out = np.fft.fftn(rho_out).ravel() / rho_in.size
        dr2 = mixer.dot(rg_out - rg_in, rg_out - rg_in)
        if it == 1 and not rediag_done:
            tighter = max(0.1 * dr2 / max(1.0, nelec), 1e-13)
            if dr2 < tol * max(1.0, nelec) and tighter < tol:
                tol = tighter
                rediag_done = True
                rho_out, occ, ef, ts = solve_all(tol)
                rg_out = np.fft.fftn(rho_out).ravel() / rho_in.size
                dr2 = mixer.dot(rg_out - rg_in, rg_out - rg_in)
        eband = float(np.sum(np.asarray(kweights)[:, None] * occ * eig))
        deband = -omega / rho_in.size * float(np.sum(rho_in * vhxc))
        etot = eband + deband + eh + exc_tot + e_ewald + ts
        hist.append({"iteration": it, "dr2": dr2, "etot": etot, "davidson_tol": tol})
        if log:
            log(f"scf {it}: E = {etot:.12f} Ry, dr2 = {dr2:.3e}, tol = {tol:.1e}")
        if dr2 < conv_thr:
            return {"converged": True, "etot": etot, "eband": eband, "deband": deband, "ehart": eh, "exc": exc_tot,
                    "ewald": e_ewald, "demet": ts, "fermi": ef, "eigenvalues": eig.copy(), "occupations": occ,
                    "rho": rho_in, "history": hist}
        rg_new = mixer.mix(rg_in, rg_out)
        rho_in = (np.fft.ifftn(rg_new.reshape(dims)) * rho_in.size).real
    return {"converged": False, "etot": hist[-1]["etot"], "eigenvalues": eig.copy(), "rho": rho_in,
            "history": hist}


def O_random(ek, nb, seed):
    """The starting block of k-point seed: the damped random block of the product's generator
    (q3x_kpoint_random_psi) is not reachable from here, so the oracle uses its own damped normal
    block; the loop's fixed point does not depend on it."""
    rng = np.random.default_rng(1000 + seed)
    z = rng.standard_normal((nb, len(ek))) + 1j * rng.standard_normal((nb, len(ek)))
    return np.asfortranarray((z / (1.0 + np.asarray(ek))[None, :]).T)
