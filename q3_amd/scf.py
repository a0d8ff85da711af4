"""Self-consistent field loop around the GPU path: the caller of H|psi> and Davidson.

Mirrors the reference's plane-wave Kohn-Sham SCF iteration (SCFIterator::iterate,
src/layer_4_computation/iterator/init.hpp:82-145; PlanewaveKohnShamSCFIterator::perform_iteration,
iterator/kohn_sham/init.hpp:357-470; iterator/kohn_sham/planewave/init.hpp:869-1091).  Per
iteration:

1. V_eff = V_loc + V_H[rho_in] + V_xc[rho_in] on the box, set on the GPU Hamiltonian;
2. every k-point solved by the GPU block Davidson, seeded by its previous eigenvectors, at the
   reference's adaptive threshold (1e-2 Ry on iteration 1, then min(tol, 0.1 dr2 / N_el),
   floor 1e-13; kohn_sham/planewave/init.hpp:869-907);
3. Fermi-Dirac occupations (iterator/internal/smearing/fermi_dirac.hpp:19-69, Fermi level by
   bisection) or fixed ones;
4. rho_out = sum_k w_k sum_n f_nk |psi_nk(r)|^2 / Omega accumulated on the GPU
   (q3x_hamiltonian_accumulate_density; charge_density/planewave/internal/accumulator.hpp);
5. dr2 = <rho_out - rho_in | rho_out - rho_in> in the Hartree metric e^2 4 pi Omega / (2 |G|^2)
   (kohn_sham/planewave/init.hpp:930-1000); converged when dr2 < conv_thr; on iteration 1 a
   re-diagonalisation at a tighter threshold when dr2 < tol N_el (kohn_sham/init.hpp:407-430);
6. Johnson modified Broyden mixing in that metric, history 8, alpha 0.7, no preconditioner
   (iterator/internal/mixing/reciprocal_space/broyden.hpp:64-380).

The host part (V_H, V_xc, mixing, occupations, energies) works on the box with numpy's FFT,
as the reference's runs it on the host with FFTW.  Energies (Ry) are the Harris-Foulkes form on
the input density, E = E_band - <rho_in|V_Hxc> + E_H + E_xc + E_Ewald - TS, equal to the
reference's total energy at self-consistency.  LDA: Perdew-Zunger 1981 (the reference's "PZ").
Ions: Gaussian-smeared point charges Z (the species' z_val, r_loc), V(G) = -4 pi Z e^2
exp(-|G|^2 r_loc^2 / 2) / (Omega |G|^2) with the finite G = 0 part 2 pi Z e^2 r_loc^2 / Omega, plus
the species' KB projectors.  The loop runs on one box: the wavefunction box must equal the density
box (the exact wavefunction grid).
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field

import numpy as np
from scipy.special import erfc

from .planewave import FP64, DeviceArray, PlanewaveBasis, PlanewaveDavidsonDiagonalizer, PlanewaveHamiltonianApplicator

E2 = 2.0  # e^2 in Rydberg units
INITIAL_TOLERANCE = 1e-2  # initial_eigenvalue_tolerance (tolerances.hpp:224)
MINIMUM_TOLERANCE = 1e-13


class BoxReciprocal:
    """|G|^2 (bohr^-2), Miller indices and the volume of a box of a crystal."""

    def __init__(self, crystal, dims):
        A = np.asarray(crystal.lattice, float).reshape(3, 3) * crystal.alat
        self.omega = abs(float(np.linalg.det(A)))
        B = 2 * math.pi * np.linalg.inv(A).T
        idx = [np.fft.fftfreq(n, 1.0 / n) for n in dims]
        m = np.stack(np.meshgrid(*idx, indexing="ij"), axis=-1)
        g = m @ B
        self.dims = tuple(dims)
        self.miller = m
        self.g2 = np.einsum("...i,...i->...", g, g)
        self.nonzero = self.g2 > 1e-12
        self.inv_g2 = np.zeros_like(self.g2)
        self.inv_g2[self.nonzero] = 1.0 / self.g2[self.nonzero]

    def to_g(self, f_r):
        return np.fft.fftn(f_r) / f_r.size

    def to_r(self, f_g):
        return np.fft.ifftn(f_g).real * f_g.size


def pz_lda(rho):
    """Perdew-Zunger 1981 LDA: energy per electron and potential, Ry."""
    rho = np.asarray(rho, float)
    e = np.zeros_like(rho)
    v = np.zeros_like(rho)
    ok = rho > 1e-30
    rs = np.cbrt(3.0 / (4.0 * math.pi * rho[ok]))
    # exchange (Hartree): -(3/4) (9 / (4 pi^2))^(1/3) / rs
    ax = -0.75 * (9.0 / (4.0 * math.pi ** 2)) ** (1.0 / 3.0)
    ex, vx = ax / rs, 4.0 / 3.0 * ax / rs
    big = rs >= 1.0
    sq = np.sqrt(np.where(big, rs, 1.0))
    den = 1.0 + 1.0529 * sq + 0.3334 * rs
    ec_hi = -0.1423 / den
    vc_hi = ec_hi * (1.0 + 7.0 / 6.0 * 1.0529 * sq + 4.0 / 3.0 * 0.3334 * rs) / den
    lnr = np.log(rs)
    ec_lo = 0.0311 * lnr - 0.048 + 0.0020 * rs * lnr - 0.0116 * rs
    vc_lo = 0.0311 * lnr - (0.048 + 0.0311 / 3.0) + 2.0 / 3.0 * 0.0020 * rs * lnr + (2.0 * -0.0116 - 0.0020) / 3.0 * rs
    e[ok] = 2.0 * (ex + np.where(big, ec_hi, ec_lo))
    v[ok] = 2.0 * (vx + np.where(big, vc_hi, vc_lo))
    return e, v


def ionic_local_potential(crystal, species, rec: BoxReciprocal):
    """Gaussian-smeared ionic potential on the box (Ry), with its finite G = 0 part."""
    vg = np.zeros(rec.dims, np.complex128)
    x = np.asarray(crystal.positions, float)
    g0 = 0.0
    for a, s_idx in enumerate(crystal.species):
        z, rl = species[s_idx]["z_val"], species[s_idx]["r_loc"]
        form = np.where(rec.nonzero, -4.0 * math.pi * z * E2 * np.exp(-rec.g2 * rl * rl / 2) * rec.inv_g2 / rec.omega,
                        0.0)
        vg += form * np.exp(-2j * math.pi * np.tensordot(rec.miller, x[a], axes=1))
        g0 += 2.0 * math.pi * z * E2 * rl * rl / rec.omega
    vg[0, 0, 0] = g0
    return rec.to_r(vg)


def ewald_energy(crystal, charges, eta=None):
    """Ewald energy (Ry) of point charges in a neutralising background (ewald_summation.hpp)."""
    A = np.asarray(crystal.lattice, float).reshape(3, 3) * crystal.alat
    omega = abs(float(np.linalg.det(A)))
    B = 2 * math.pi * np.linalg.inv(A).T
    tau = np.asarray(crystal.positions, float) @ A
    z = np.asarray(charges, float)
    eta = math.sqrt(math.pi) / omega ** (1.0 / 3.0) if eta is None else eta
    reach = 6.0 / eta  # erfc(6) < 3e-17
    heights = [omega / np.linalg.norm(np.cross(A[(i + 1) % 3], A[(i + 2) % 3])) for i in range(3)]
    n = [int(reach / h) + 2 for h in heights]
    cells = np.stack(np.meshgrid(*[np.arange(-k, k + 1) for k in n], indexing="ij"), -1).reshape(-1, 3) @ A
    d = tau[:, None, None, :] - tau[None, :, None, :] + cells[None, None, :, :]
    r = np.linalg.norm(d, axis=-1)
    zz = (z[:, None] * z[None, :])[:, :, None] * np.ones_like(r)
    keep = r > 1e-10
    e_real = 0.5 * float(np.sum(zz[keep] * erfc(eta * r[keep]) / r[keep]))
    gcut = 12.0 * eta  # exp(-G^2 / 4 eta^2) < 1e-15
    m = [int(gcut / np.linalg.norm(B[i])) + 2 for i in range(3)]
    mm = np.stack(np.meshgrid(*[np.arange(-k, k + 1) for k in m], indexing="ij"), -1).reshape(-1, 3)
    mm = mm[np.any(mm != 0, axis=1)]
    G = mm @ B
    g2 = np.einsum("ij,ij->i", G, G)
    sf = np.exp(1j * (G @ tau.T)) @ z
    e_rec = 2.0 * math.pi / omega * float(np.sum(np.exp(-g2 / (4.0 * eta * eta)) / g2 * np.abs(sf) ** 2))
    e_self = -eta / math.sqrt(math.pi) * float(z @ z)
    e_bg = -math.pi * float(z.sum()) ** 2 / (2.0 * omega * eta * eta)
    return E2 * (e_real + e_rec + e_self + e_bg)


def fermi_dirac(eig, kweights, nelec, sigma):
    """Occupations (electrons per band, two per band at most), Fermi level and -TS (Ry)."""
    eig = np.asarray(eig, float)
    w = np.asarray(kweights, float)[:, None]
    if sigma <= 0.0:
        occ = np.zeros_like(eig)
        occ[:, : nelec // 2] = 2.0
        return occ, float(eig[:, : nelec // 2].max()), 0.0
    lo, hi = float(eig.min()) - 10.0 * sigma, float(eig.max()) + 10.0 * sigma
    for _ in range(200):  # bisection on the electron count
        ef = 0.5 * (lo + hi)
        with np.errstate(over="ignore"):
            n = float(np.sum(w * 2.0 / (np.exp((eig - ef) / sigma) + 1.0)))
        lo, hi = (ef, hi) if n < nelec else (lo, ef)
        if hi - lo < 1e-15:
            break
    ef = 0.5 * (lo + hi)
    with np.errstate(over="ignore"):
        f = 1.0 / (np.exp((eig - ef) / sigma) + 1.0)
    f_log = np.where(f > 1e-16, f * np.log(np.maximum(f, 1e-300)), 0.0)
    g_log = np.where(1.0 - f > 1e-16, (1.0 - f) * np.log(np.maximum(1.0 - f, 1e-300)), 0.0)
    return 2.0 * f, ef, float(np.sum(w * 2.0 * sigma * (f_log + g_log)))


class BroydenMixer:
    """Johnson's modified Broyden on rho(G) with metric weights w (broyden.hpp:64-380): secant
    pairs (dF = F_prev - F, drho = rho_prev - rho) in a ring of depth entries, the correction
    gamma = beta^-1 <dF|F>, then rho + alpha F."""

    def __init__(self, weights, alpha=0.7, depth=8):
        self.w = weights.ravel()
        self.alpha, self.depth = alpha, depth
        self.count = 0
        self.pairs = []
        self.last = None

    def inner(self, a, b):
        return float(np.dot(self.w, (a.conj() * b).real))

    def mix(self, rho_in, rho_out):
        self.count += 1
        x = rho_in.ravel().copy()
        F = rho_out.ravel() - x
        if self.last is not None:
            pair = (self.last[0] - F, self.last[1] - x)
            if len(self.pairs) < self.depth:
                self.pairs.append(pair)
            else:
                self.pairs[(self.count - 2) % self.depth] = pair
        self.last = (F.copy(), x.copy())
        if self.pairs:
            beta = np.array([[self.inner(p[0], q[0]) for q in self.pairs] for p in self.pairs])
            gamma = np.linalg.solve(beta, np.array([self.inner(p[0], F) for p in self.pairs]))
            for g, (df, dr) in zip(gamma, self.pairs):
                x -= g * dr
                F -= g * df
        return (x + self.alpha * F).reshape(rho_in.shape)


@dataclass
class SCFResult:
    converged: bool
    iterations: int
    etot: float
    parts: dict
    eigenvalues: np.ndarray
    occupations: np.ndarray
    fermi: float
    rho: np.ndarray
    history: list = field(default_factory=list)


class PlanewaveSCF:
    """The SCF loop on one GPU: Davidson and the density on the device, the potential, mixing and
    occupations on the host."""

    def __init__(self, context, crystal, species, ecutwfc_ry, kpoints, kweights, nbands, nelec,
                 fft_factors=(1, 1, 1), sigma=0.01, alpha=0.7, history=8, conv_thr=1e-10, max_iter=60,
                 davidson_max_iter=40, precision=FP64, hold_every_band=True):
        self.ctx = context
        self.crystal, self.species = crystal, species
        self.basis = PlanewaveBasis(crystal, ecutwfc_ry, 0.0, True, fft_factors)
        if tuple(self.basis.wave_dims) != tuple(self.basis.rho_dims):
            raise ValueError("the SCF loop runs on one box: the wavefunction box must equal the density box")
        self.dims = tuple(self.basis.wave_dims)
        self.rec = BoxReciprocal(crystal, self.dims)
        self.kpoints = np.asarray(kpoints, float).reshape(-1, 3)
        self.kweights = np.asarray(kweights, float)
        self.nbands, self.nelec = nbands, nelec
        self.sigma, self.conv_thr, self.max_iter = sigma, conv_thr, max_iter
        self.davidson_max_iter, self.precision = davidson_max_iter, precision
        self.n_converged = nbands if hold_every_band else nelec // 2
        self.metric = np.where(self.rec.nonzero, E2 * 4.0 * math.pi * self.rec.omega * 0.5 * self.rec.inv_g2, 0.0)
        self.mixer = BroydenMixer(self.metric, alpha, history)
        self.vloc = ionic_local_potential(crystal, species, self.rec)
        self.ewald = ewald_energy(crystal, [species[s]["z_val"] for s in crystal.species])
        from . import workloads as W
        self.bs, self.db = W.d_blocks_for(crystal, species)
        self.kps = [self.basis.compute_gk(k) for k in self.kpoints]
        self.betas = [kp.kb_projectors(crystal, species) for kp in self.kps]
        self.h = PlanewaveHamiltonianApplicator(context, self.basis)
        self.psi = [kp.random_psi(nbands, i) for i, kp in enumerate(self.kps)]
        npw_max = max(kp.npw for kp in self.kps)
        self.d_psi = DeviceArray(context, npw_max * nbands * 16)
        self.d_rho = DeviceArray(context, int(np.prod(self.dims)) * 8)
        self.solves = 0
        self.davidson_iterations = 0

    def _potential(self, rho):
        rg = self.rec.to_g(rho)
        vh = self.rec.to_r(4.0 * math.pi * E2 * rg * self.rec.inv_g2)
        eh = 0.5 * self.rec.omega * float(np.sum(4.0 * math.pi * E2 * np.abs(rg) ** 2 * self.rec.inv_g2))
        exc, vxc = pz_lda(rho)
        return vh + vxc, eh, self.rec.omega / rho.size * float(np.sum(exc * rho))

    def _solve_all(self, tol):
        eig = np.zeros((len(self.kps), self.nbands))
        for i, kp in enumerate(self.kps):
            self.h.prepare_kpoint(kp)
            self.h.set_projectors(self.betas[i], self.bs, self.db)
            d = PlanewaveDavidsonDiagonalizer(self.h)
            d.set_tolerance(tol)
            d.set_maximum_iterations(self.davidson_max_iter)
            eig[i], self.psi[i] = d.solve(self.nbands, self.n_converged, self.psi[i], precision=self.precision)
            self.solves += 1
            self.davidson_iterations += d.report.iterations
        occ, ef, ts = fermi_dirac(eig, self.kweights, self.nelec, self.sigma)
        self.d_rho.upload(np.zeros(int(np.prod(self.dims))))
        for i, kp in enumerate(self.kps):
            self.h.prepare_kpoint(kp)
            self.d_psi.upload(self.psi[i])
            self.h.accumulate_density_device(self.d_psi.ptr, self.nbands, self.kweights[i] * occ[i] / self.rec.omega,
                                             self.d_rho.ptr)
        self.ctx.synchronize()
        rho = self.d_rho.download((int(np.prod(self.dims)),), np.float64).reshape(self.dims)
        return rho, eig, occ, ef, ts

    def run(self, log=None) -> SCFResult:
        rho_in = np.full(self.dims, self.nelec / self.rec.omega)
        tol = INITIAL_TOLERANCE
        hist = []
        rediag_done = False
        res = None
        for it in range(1, self.max_iter + 1):
            vhxc, eh, exc = self._potential(rho_in)
            self.h.update_potential(self.vloc + vhxc)
            if it > 1:
                tol = max(min(tol, 0.1 * hist[-1]["dr2"] / max(1.0, self.nelec)), MINIMUM_TOLERANCE)
            rho_out, eig, occ, ef, ts = self._solve_all(tol)
            rg_in, rg_out = self.rec.to_g(rho_in), self.rec.to_g(rho_out)
            dr2 = self.mixer.inner((rg_out - rg_in).ravel(), (rg_out - rg_in).ravel())
            if it == 1 and not rediag_done:
                tighter = max(0.1 * dr2 / max(1.0, self.nelec), MINIMUM_TOLERANCE)
                if dr2 < tol * max(1.0, self.nelec) and tighter < tol:
                    tol, rediag_done = tighter, True
                    rho_out, eig, occ, ef, ts = self._solve_all(tol)
                    rg_out = self.rec.to_g(rho_out)
                    dr2 = self.mixer.inner((rg_out - rg_in).ravel(), (rg_out - rg_in).ravel())
            eband = float(np.sum(self.kweights[:, None] * occ * eig))
            deband = -self.rec.omega / rho_in.size * float(np.sum(rho_in * vhxc))
            parts = {"eband": eband, "deband": deband, "ehart": eh, "exc": exc, "ewald": self.ewald, "demet": ts}
            etot = sum(parts.values())
            hist.append({"iteration": it, "dr2": dr2, "etot": etot, "davidson_tol": tol})
            if log:
                log(f"scf {it}: E = {etot:.12f} Ry, dr2 = {dr2:.3e}, Davidson tol = {tol:.1e}")
            res = SCFResult(dr2 < self.conv_thr, it, etot, parts, eig, occ, ef, rho_in, hist)
            if res.converged:
                return res
            rho_in = self.rec.to_r(self.mixer.mix(rg_in, rg_out))
        return res

    def close(self):
        for a in (getattr(self, "d_psi", None), getattr(self, "d_rho", None)):
            if a is not None:
                a.free()
        if getattr(self, "h", None) is not None:
            self.h.close()
            self.h = None


__all__ = ["PlanewaveSCF", "SCFResult", "pz_lda", "ewald_energy", "fermi_dirac", "BroydenMixer",
           "ionic_local_potential", "BoxReciprocal"]
