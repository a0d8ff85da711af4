"""Synthetic inputs of the path: crystals, the synthetic norm-conserving pseudopotential,
the local potential V(r), KB D blocks and starting wavefunctions, for the five workloads

  cfg1  Si2 FCC a=10.26, ecutwfc 15 Ha, FFT 30^3, 4x4x4 shifted MP, 8 bands (CPU-runnable)
  cfg2  Si64 (2x2x2 conventional, a=20.52, displaced N(0,0.02) seed 1), Gamma, ecutwfc 25 Ha,
        FFT 96^3, 160 bands (128 occupied), 256 projectors
  cfg3  cfg2 in mixed precision
  cfg4  Si8 cubic a=10.26, ecutwfc 20 Ha, FFT 48^3, 48 bands, 64 random k (seed 3)
  cfg5  serving stream of band-structure requests (seed 4)

All energies in Ry inside (1 Ha = 2 Ry).  Nothing here runs on the GPU.

Synthetic pseudopotential (seed 0), Z_val = 4:
  radial mesh   r_i = exp(-7 + 0.0125 i) / 14, i = 0..1084, rab_i = 0.0125 r_i
  projectors    s: beta(r) = exp(-r^2 / (2 * 0.5^2)),  p: beta(r) = r exp(-r^2 / (2 * 0.6^2)),
                stored as r*beta(r) (UPF convention), cutoff index = #points with r <= 10 bohr
  D_ion         diag(D_s, D_p), D ~ U[1, 5] Ha from numpy default_rng(0), stored in Ry
  local part    v(G) = -4 pi Z exp(-G^2 r_loc^2 / 2) / (Omega G^2), r_loc = 0.44, v(0) = 0;
                V(r) = sum_G sum_atoms v(G) exp(-i G.tau) exp(i G.r) over the density sphere
"""
from __future__ import annotations

import math
from dataclasses import dataclass

import numpy as np

from .planewave import Crystal

HA = 2.0  # Ry per Ha

FCC = np.array([[-0.5, 0.0, 0.5], [0.0, 0.5, 0.5], [-0.5, 0.5, 0.0]])  # ibrav 2 (lattice.hpp)
SC = np.eye(3)
DIAMOND_CONVENTIONAL = np.array([[0, 0, 0], [0, .5, .5], [.5, 0, .5], [.5, .5, 0],
                                 [.25, .25, .25], [.25, .75, .75], [.75, .25, .75], [.75, .75, .25]])

Z_VAL = 4.0
R_LOC = 0.44
SIGMA_S = 0.5
SIGMA_P = 0.6


def synthetic_species(seed: int = 0):
    """The synthetic norm-conserving Si-like species (dict consumed by KPoint.kb_projectors)."""
    n = 1085
    dx = 0.0125
    r = np.exp(-7.0 + dx * np.arange(n)) / 14.0
    rab = dx * r
    cutoff = int(np.count_nonzero(r <= 10.0))
    beta_s = r * np.exp(-r * r / (2 * SIGMA_S ** 2))            # r * beta_s
    beta_p = r * (r * np.exp(-r * r / (2 * SIGMA_P ** 2)))      # r * beta_p
    rng = np.random.default_rng(seed)
    d_ha = rng.uniform(1.0, 5.0, size=2)
    return {"r": r, "rab": rab, "l": [0, 1], "cutoff": [cutoff, cutoff],
            "beta_r": np.vstack([beta_s, beta_p]), "dion_ry": np.diag(d_ha * HA), "z_val": Z_VAL,
            "r_loc": R_LOC}


def write_synthetic_upf(path: str, seed: int = 0) -> None:
    """The synthetic species of a seed written as a UPF version 2 file (q3_amd.upf); reading it
    back with upf.read_upf gives the same arrays bit for bit."""
    from . import upf
    upf.write_upf_v2(path, synthetic_species(seed), element="Si",
                     info=f"Synthetic norm-conserving Si-like species, seed {seed}: Z_val 4, Gaussian local "
                          f"part r_loc {R_LOC}, s and p projectors r^l exp(-r^2 / 2 sigma^2) with sigma "
                          f"{SIGMA_S} and {SIGMA_P} bohr, D ~ U[1, 5] Ha.")


def d_blocks_for(crystal: Crystal, species_list):
    """Per-atom D blocks (column-major, packed) and their sizes, built by the library
    (q3x_kb_d_blocks): D_ij = dion[beta_i][beta_j] when (l,m) of the two channels agree
    (norm_conserving/init.hpp:84-100)."""
    import ctypes

    from ._lib import check, lib
    asp = np.ascontiguousarray(crystal.species, dtype=np.int32)
    nbeta = np.ascontiguousarray([len(s["l"]) for s in species_list], dtype=np.int32)
    l_all = np.ascontiguousarray(np.concatenate([np.asarray(s["l"], np.int32) for s in species_list]), np.int32)
    dion = np.ascontiguousarray(np.concatenate([np.asarray(s["dion_ry"], np.float64).ravel()
                                                for s in species_list]), np.float64)
    n = ctypes.c_long()
    p = lambda a: a.ctypes.data_as(ctypes.c_void_p)  # noqa: E731
    check(lib().q3x_kb_d_blocks(len(asp), p(asp), len(species_list), p(nbeta), p(l_all), p(dion), None, None,
                                ctypes.byref(n)))
    sizes = np.zeros(len(asp), np.int32)
    blocks = np.zeros(max(n.value, 1))
    check(lib().q3x_kb_d_blocks(len(asp), p(asp), len(species_list), p(nbeta), p(l_all), p(dion), p(sizes),
                                p(blocks), ctypes.byref(n)))
    return sizes, blocks[:n.value]


def local_potential(crystal: Crystal, basis, species_list, grid: str = "wave") -> np.ndarray:
    """V(r) on the wave grid (grid="wave") or the density grid (grid="rho"), C order (i1,i2,i3),
    Ry.  Built in G space over the density sphere and brought to real space with the
    FFTW_BACKWARD convention (numpy: N * ifftn)."""
    m, g, g2 = basis.g_vectors()
    if getattr(basis, "gamma_only", False):  # half sphere: add the mirror of every G but G = 0
        nz0 = np.any(m != 0, axis=1)
        m, g, g2 = np.vstack([m, -m[nz0]]), np.vstack([g, -g[nz0]]), np.concatenate([g2, g2[nz0]])
    tpa = 2 * math.pi / crystal.alat
    gabs2 = g2 * tpa * tpa  # bohr^-2
    omega = basis.cell_volume
    vg = np.zeros(len(g2), np.complex128)
    nz = gabs2 > 1e-12
    pos = np.asarray(crystal.positions, float)
    # structure factors exp(-i 2 pi m . x) (crystal coordinates x, Miller m; = exp(-i G . tau)) of
    # each species over the Miller box that holds the sphere, as products of per-axis phases
    # summed over the atoms by one matrix product: O(atoms x box) multiply-adds, not exps
    half = np.abs(m).max(axis=0)
    axes = [np.arange(-h, h + 1) for h in half]
    for s_idx, sp in enumerate(species_list):
        form = np.zeros(len(g2))
        form[nz] = -4 * math.pi * sp["z_val"] * np.exp(-gabs2[nz] * sp["r_loc"] ** 2 / 2) / (omega * gabs2[nz])
        x = pos[np.asarray(crystal.species) == s_idx]
        if len(x) == 0:
            continue
        e = [np.exp(-2j * math.pi * np.outer(x[:, d], axes[d])) for d in range(3)]  # atoms x n_d
        p12 = (e[0][:, :, None] * e[1][:, None, :]).reshape(len(x), -1)
        sbox = (p12.T @ e[2]).reshape(len(axes[0]), len(axes[1]), len(axes[2]))
        sf = sbox[m[:, 0] + half[0], m[:, 1] + half[1], m[:, 2] + half[2]]
        vg += form * sf
    dims = basis.wave_dims if grid == "wave" else basis.rho_dims
    box = np.zeros(dims, np.complex128)
    ok = np.all(np.abs(m) <= np.array(dims) // 2, axis=1)
    w = np.where(m < 0, m + np.array(dims), m)
    box[w[ok, 0], w[ok, 1], w[ok, 2]] = vg[ok]
    v = np.fft.ifftn(box) * box.size
    return np.ascontiguousarray(v.real)


def random_psi(kinetic_ry: np.ndarray, nb: int, seed: int) -> np.ndarray:
    """Complex normal coefficients damped by 1/(1+|k+G|^2), npw x nb (Fortran order)."""
    rng = np.random.default_rng(seed)
    npw = len(kinetic_ry)
    z = rng.standard_normal((nb, npw)) + 1j * rng.standard_normal((nb, npw))
    z = z / (1.0 + kinetic_ry)[None, :]
    return np.asfortranarray(z.T)


@dataclass
class Workload:
    name: str
    crystal: Crystal
    ecutwfc_ry: float
    fft_factors: tuple
    kpoints: np.ndarray
    nbands: int
    noccupied: int
    species: list


def cfg1() -> Workload:
    c = Crystal(FCC.copy(), 10.26, np.array([[0, 0, 0], [0.25, 0.25, 0.25]]), np.zeros(2, np.int32))
    g = (np.arange(4) + 0.5) / 4.0  # 4x4x4 shifted Monkhorst-Pack, crystal coordinates
    ks = np.array([[a, b, c_] for a in g for b in g for c_ in g])
    ks = ks - np.round(ks)
    return Workload("cfg1_si2_fcc_15ha", c, 15.0 * HA, (2, 2, 2), ks, 8, 4, [synthetic_species(0)])


def si64_positions(seed: int = 1, sigma_bohr: float = 0.02, alat: float = 20.52):
    cells = np.array([[i, j, k] for i in range(2) for j in range(2) for k in range(2)])
    frac = np.array([(c + b) / 2.0 for c in cells for b in DIAMOND_CONVENTIONAL])
    rng = np.random.default_rng(seed)
    frac = frac + rng.normal(0.0, sigma_bohr, size=frac.shape) / alat
    return frac


def cfg2() -> Workload:
    c = Crystal(SC.copy(), 20.52, si64_positions(1), np.zeros(64, np.int32))
    return Workload("cfg2_si64_gamma_25ha", c, 25.0 * HA, (1, 1, 1), np.zeros((1, 3)), 160, 128,
                    [synthetic_species(0)])


def cfg4() -> Workload:
    c = Crystal(SC.copy(), 10.26, DIAMOND_CONVENTIONAL.astype(float), np.zeros(8, np.int32))
    rng = np.random.default_rng(3)
    ks = rng.uniform(-0.5, 0.5, size=(64, 3))
    return Workload("cfg4_si8_64k_20ha", c, 20.0 * HA, (4, 4, 4), ks, 48, 16, [synthetic_species(0)])


def small_test_cell(ecutwfc_ry: float = 12.0) -> Workload:
    """Tiny Si2 FCC cell for fast parity tests."""
    w = cfg1()
    w.ecutwfc_ry = ecutwfc_ry
    w.name = "si2_fcc_small"
    return w


# ---------------------------------------------------------------- cfg5: serving stream ----
CARBON_R_LOC = 0.35
# (lattice kind, supercell repetitions of the conventional cube, atoms): primitive FCC cell to
# the 3x3x3 conventional supercell
SUPERCELLS = [("fcc", (1, 1, 1), 2), ("sc", (1, 1, 1), 8), ("sc", (2, 1, 1), 16), ("sc", (2, 2, 1), 32),
              ("sc", (2, 2, 2), 64), ("sc", (3, 2, 2), 96), ("sc", (3, 3, 2), 144), ("sc", (3, 3, 3), 216)]
A_SI, A_C = 10.26, 6.74


def carbon_like_species():
    """Second synthetic species of the Si/C alloy: same form, D from seed 1, r_loc 0.35 bohr."""
    sp = synthetic_species(1)
    sp["r_loc"] = CARBON_R_LOC
    return sp


@dataclass
class ServingRequest:
    arrival_s: float
    crystal: Crystal
    ecutwfc_ry: float
    kpoints: np.ndarray
    nbands: int
    noccupied: int
    carbon_fraction: float


def alloy_cell(reps, kind: str, x: float, rng) -> Crystal:
    """Si_(1-x) C_x diamond cell: each site is C with probability x; a interpolated linearly."""
    a = A_SI + (A_C - A_SI) * x
    if kind == "fcc":
        lattice, frac = FCC.copy(), np.array([[0.0, 0.0, 0.0], [0.25, 0.25, 0.25]])
    else:
        r = np.array(reps, dtype=float)
        cells = np.array([[i, j, k] for i in range(reps[0]) for j in range(reps[1]) for k in range(reps[2])])
        frac = np.array([(c + b) / r for c in cells for b in DIAMOND_CONVENTIONAL])
        lattice = np.diag(r)
    species = (rng.uniform(size=len(frac)) < x).astype(np.int32)
    return Crystal(lattice, a, frac, species)


def cfg5_requests(n_requests: int = 2000, rate_per_s: float = 20.0, seed: int = 4,
                  max_atoms: int = 216) -> list:
    """Poisson stream of band-structure requests (seed 4): Si/C alloy cells from the primitive
    cell to 3x3x3 conventional (C fraction U[0, 0.5], a from 10.26 to 6.74 bohr), ecutwfc
    U[15, 25] Ha, 8-40 k-points along a random straight line, bands = 2 N_atoms + 8.
    max_atoms < 216 draws the cell among the smaller supercells only."""
    rng = np.random.default_rng(seed)
    cells = [c for c in SUPERCELLS if c[2] <= max_atoms]
    t = 0.0
    out = []
    for _ in range(n_requests):
        t += rng.exponential(1.0 / rate_per_s)
        kind, reps, nat = cells[int(rng.integers(len(cells)))]
        x = float(rng.uniform(0.0, 0.5))
        crystal = alloy_cell(reps, kind, x, rng)
        ecut = float(rng.uniform(15.0, 25.0)) * HA
        nk = int(rng.integers(8, 41))
        k0, k1 = rng.uniform(-0.5, 0.5, 3), rng.uniform(-0.5, 0.5, 3)
        ks = k0[None, :] + np.linspace(0.0, 1.0, nk)[:, None] * (k1 - k0)[None, :]
        out.append(ServingRequest(t, crystal, ecut, ks, 2 * nat + 8, 2 * nat, x))
    return out


def serving_species():
    return [synthetic_species(0), carbon_like_species()]
