"""ctypes front end of the FP64 CPU oracle (oracle/q3_oracle.cpp).

Test infrastructure only: tests/, ``__graft_entry__.smoke()`` and bench.py's
cpu_baseline leg call this module as the checker.  The product library never
imports it.  Builds with ``make -C oracle`` (``build()`` in __graft_entry__ does it).
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "_build", "libq3oracle.so")
_lib = None

_i = ctypes.c_int
_d = ctypes.c_double
_p = ctypes.c_void_p


def build() -> str:
    subprocess.run(["make", "-s", "-C", _HERE], check=True)
    # the reference's own footprint estimate, compiled from its tree when that is present
    subprocess.run(["make", "-s", "-C", _HERE, "ref"], check=True)
    return _LIB_PATH


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = ctypes.CDLL(_LIB_PATH)
        L.orc_good_fft_order.restype = _i
        L.orc_good_fft_order.argtypes = [_i, _i]
        L.orc_crystal.argtypes = [_p, _d, _p, _p, _p]
        L.orc_grid_dims.argtypes = [_p, _d, _d, _p, _p]
        L.orc_basis_create.restype = _p
        L.orc_basis_create.argtypes = [_p, _d, _d, _d, _i, _p]
        L.orc_basis_destroy.argtypes = [_p]
        L.orc_basis_info.argtypes = [_p, _p, _p, _p]
        L.orc_basis_gvectors.argtypes = [_p, _p, _p, _p]
        L.orc_kpoint.restype = _i
        L.orc_kpoint.argtypes = [_p, _p, _p, _p, _p, _i]
        L.orc_fft3d.argtypes = [_p, _p, _i]
        L.orc_hamiltonian_apply.argtypes = [_p, _i, _p, _p, _p, _i, _p, _i, _p, _p, _i, _p, _p]
        L.orc_beta_table_size.restype = _i
        L.orc_beta_table_size.argtypes = [_d]
        L.orc_beta_table.argtypes = [_i, _p, _p, _i, _i, _p, _i, _p]
        L.orc_sph_bessel.restype = _d
        L.orc_sph_bessel.argtypes = [_i, _d]
        L.orc_ylm.argtypes = [_i, _i, _p, _p]
        L.orc_lagrange4.restype = _d
        L.orc_lagrange4.argtypes = [_p, _i, _d, _d]
        L.orc_simpson.restype = _d
        L.orc_simpson.argtypes = [_p, _p, _i]
        L.orc_sph_bessel_batch.argtypes = [_i, _d, _p, _i, _p]
        L.orc_kb_projectors.restype = _i
        L.orc_kb_projectors.argtypes = [_p, _p, _i, _p, _i, _p, _p, _i, _p, _p, _p, _p, _p, _p, _p, _p]
        L.orc_restrict_potential.argtypes = [_p, _p, _p, _p]
        L.orc_d_blocks.restype = ctypes.c_long
        L.orc_d_blocks.argtypes = [_i, _p, _i, _p, _p, _p, _p, _p]
        L.orc_zhegv.restype = _i
        L.orc_zhegv.argtypes = [_i, _p, _p, _p, _p]
        L.orc_davidson.restype = _i
        L.orc_davidson.argtypes = [_p, _i, _p, _p, _p, _i, _p, _i, _p, _p, _i, _i, _p, _i, _p, _d,
                                   _i, _i, _p]
        L.orc_davidson_stage_ms.argtypes = [_p]
        L.orc_basis_create2.restype = _p
        L.orc_basis_create2.argtypes = [_p, _d, _d, _d, _i, _p, _i]
        L.orc_kpoint_minus_map.restype = _i
        L.orc_kpoint_minus_map.argtypes = [_p, _p, _p, _i]
        L.orc_hamiltonian_apply2.argtypes = [_p, _i, _p, _p, _p, _p, _i, _p, _i, _p, _p, _i, _p, _p]
        L.orc_davidson2.restype = _i
        L.orc_davidson2.argtypes = [_p, _i, _p, _p, _p, _p, _i, _p, _i, _p, _p, _i, _i, _p, _i, _p, _d, _i, _i, _p]
        L.orc_set_threads.argtypes = [_i]
        L.orc_get_threads.restype = _i
        _lib = L
    return _lib


def _ptr(a: np.ndarray):
    return a.ctypes.data_as(ctypes.c_void_p)


def _c(a, dtype):
    return np.ascontiguousarray(a, dtype=dtype)


def good_fft_order(n: int, factor: int = 1) -> int:
    return lib().orc_good_fft_order(n, factor)


def crystal(lat, alat):
    lat = _c(lat, np.float64)
    recip = np.zeros(9)
    vol = np.zeros(1)
    tpa2 = np.zeros(1)
    lib().orc_crystal(_ptr(lat), alat, _ptr(recip), _ptr(vol), _ptr(tpa2))
    return recip.reshape(3, 3), float(vol[0]), float(tpa2[0])


def grid_dims(lat, alat, ecut, factors=(1, 1, 1)):
    lat = _c(lat, np.float64)
    f = _c(factors, np.int32)
    dims = np.zeros(3, np.int32)
    lib().orc_grid_dims(_ptr(lat), alat, ecut, _ptr(f), _ptr(dims))
    return tuple(int(x) for x in dims)


class Basis:
    """G sphere and FFT grids (PlanewaveBasis::generate)."""

    def __init__(self, lat, alat, ecutwfc, ecutrho=0.0, use_exact=True, factors=(1, 1, 1), gamma_only=False):
        self._lat = _c(lat, np.float64)
        f = _c(factors, np.int32)
        self.gamma_only = bool(gamma_only)
        self.h = lib().orc_basis_create2(_ptr(self._lat), alat, ecutwfc, ecutrho, int(use_exact), _ptr(f),
                                         int(gamma_only))
        rho = np.zeros(3, np.int32)
        wave = np.zeros(3, np.int32)
        ng = np.zeros(1, np.int32)
        lib().orc_basis_info(self.h, _ptr(rho), _ptr(wave), _ptr(ng))
        self.rho_dims = tuple(int(x) for x in rho)
        self.wave_dims = tuple(int(x) for x in wave)
        self.ng = int(ng[0])

    def gvectors(self):
        m = np.zeros((self.ng, 3), np.int32)
        g = np.zeros((self.ng, 3))
        g2 = np.zeros(self.ng)
        lib().orc_basis_gvectors(self.h, _ptr(m), _ptr(g), _ptr(g2))
        return m, g, g2

    def kpoint(self, k):
        k = _c(k, np.float64)
        cap = self.ng
        gidx = np.zeros(cap, np.int32)
        ek = np.zeros(cap)
        fm = np.zeros(cap, np.int32)
        n = lib().orc_kpoint(self.h, _ptr(k), _ptr(gidx), _ptr(ek), _ptr(fm), cap)
        return gidx[:n].copy(), ek[:n].copy(), fm[:n].copy()

    def minus_map(self, k):
        """Box offsets of -G for the k-point's plane waves (the Gamma trick's mirror map)."""
        k = _c(k, np.float64)
        mm = np.zeros(self.ng, np.int32)
        n = lib().orc_kpoint_minus_map(self.h, _ptr(k), _ptr(mm), self.ng)
        return mm[:n].copy()

    def kb_projectors(self, k, gidx, atom_pos, atom_species, species):
        """species: list of dicts with r, rab, l (list), cutoff (list), beta_r (nbeta x nmesh)."""
        k = _c(k, np.float64)
        gidx = _c(gidx, np.int32)
        pos = _c(atom_pos, np.float64)
        asp = _c(atom_species, np.int32)
        nmesh = _c([len(s["r"]) for s in species], np.int32)
        r_all = _c(np.concatenate([s["r"] for s in species]), np.float64)
        rab_all = _c(np.concatenate([s["rab"] for s in species]), np.float64)
        nbeta = _c([len(s["l"]) for s in species], np.int32)
        l_all = _c(np.concatenate([s["l"] for s in species]), np.int32)
        cut_all = _c(np.concatenate([s["cutoff"] for s in species]), np.int32)
        br_all = _c(np.concatenate([np.asarray(s["beta_r"]).ravel() for s in species]), np.float64)
        nproj = sum(sum(2 * l + 1 for l in species[s]["l"]) for s in asp)
        out = np.zeros((nproj, len(gidx)), np.complex128)
        n = lib().orc_kb_projectors(self.h, _ptr(k), len(gidx), _ptr(gidx), len(asp), _ptr(pos), _ptr(asp),
                                    len(species), _ptr(nmesh), _ptr(r_all), _ptr(rab_all), _ptr(nbeta),
                                    _ptr(l_all), _ptr(cut_all), _ptr(br_all), _ptr(out))
        assert n == nproj
        return out.T.copy(order="F")  # npw x nproj, column-major

    def __del__(self):
        if getattr(self, "h", None):
            lib().orc_basis_destroy(self.h)
            self.h = None


def fft3d(data: np.ndarray, sign: int) -> np.ndarray:
    a = _c(data, np.complex128).copy()
    dims = _c(a.shape, np.int32)
    lib().orc_fft3d(_ptr(dims), _ptr(a), sign)
    return a


def beta_table(r, rab, l, cutoff, beta_r, ecutrho):
    nq = lib().orc_beta_table_size(ecutrho)
    out = np.zeros(nq)
    r = _c(r, np.float64)
    rab = _c(rab, np.float64)
    br = _c(beta_r, np.float64)
    lib().orc_beta_table(len(r), _ptr(r), _ptr(rab), l, cutoff, _ptr(br), nq, _ptr(out))
    return out


def sph_bessel(l, x):
    return lib().orc_sph_bessel(l, x)


def sph_bessel_batch(l, q, r):
    r = _c(r, np.float64)
    out = np.zeros(len(r))
    lib().orc_sph_bessel_batch(l, q, _ptr(r), len(r), _ptr(out))
    return out


def lagrange4(table, q, dq):
    """4-point Lagrange interpolation of a uniform table (lagrange.hpp:16-30)."""
    t = _c(table, np.float64)
    return lib().orc_lagrange4(_ptr(t), len(t), q, dq)


def simpson(f, rab):
    """Simpson's rule on a radial mesh with derivative rab (simpson.hpp:22-48)."""
    f = _c(f, np.float64)
    rab = _c(rab, np.float64)
    return lib().orc_simpson(_ptr(f), _ptr(rab), len(f))


def ylm(lmax, g):
    g = _c(g, np.float64)
    n = g.shape[0]
    out = np.zeros((n, (lmax + 1) ** 2))
    lib().orc_ylm(lmax, n, _ptr(g), _ptr(out))
    return out


def _ham_args(dims, fftmap, ekin, V, beta, block_sizes, dblocks):
    dims = _c(dims, np.int32)
    fm = _c(fftmap, np.int32)
    ek = _c(ekin, np.float64)
    Vr = _c(V, np.float64).ravel()
    if beta is None or beta.size == 0:
        b = np.zeros(1, np.complex128)
        nproj = 0
    else:
        b = np.asfortranarray(beta, dtype=np.complex128)
        nproj = beta.shape[1]
    bs = _c(block_sizes if block_sizes is not None else [], np.int32)
    db = _c(dblocks if dblocks is not None else [0.0], np.float64)
    return dims, fm, ek, Vr, b, nproj, bs, db


def hamiltonian_apply(dims, fftmap, ekin, V, beta, block_sizes, dblocks, psi, minus_map=None):
    """psi: npw x nvec (column-major); returns H psi of the same shape.  With minus_map (the -G box
    offsets of a Gamma-trick basis) psi holds half-sphere vectors of real functions."""
    dims, fm, ek, Vr, b, nproj, bs, db = _ham_args(dims, fftmap, ekin, V, beta, block_sizes, dblocks)
    psi = np.asfortranarray(psi, dtype=np.complex128)
    npw, nvec = psi.shape
    out = np.zeros_like(psi, order="F")
    mm = None if minus_map is None else _c(minus_map, np.int32)
    lib().orc_hamiltonian_apply2(_ptr(dims), npw, _ptr(fm), None if mm is None else _ptr(mm), _ptr(ek), _ptr(Vr),
                                 nproj, _ptr(b), len(bs), _ptr(bs), _ptr(db), nvec, _ptr(psi), _ptr(out))
    return out


def restrict_potential(v_dense, smooth_dims):
    """V on the density grid (C order) restricted to the smooth box in G space
    (planewave_potential_utils.hpp:33-130)."""
    v = _c(v_dense, np.float64)
    dd = _c(v.shape, np.int32)
    sd = _c(smooth_dims, np.int32)
    out = np.zeros(tuple(int(x) for x in smooth_dims))
    lib().orc_restrict_potential(_ptr(dd), _ptr(v), _ptr(sd), _ptr(out))
    return out


def d_blocks(atom_species, species):
    """Per-atom D blocks (sizes, packed column-major blocks) of the nonlocal term
    (d_matrix_element, norm_conserving/init.hpp:84-100)."""
    asp = _c(atom_species, np.int32)
    nbeta = _c([len(s["l"]) for s in species], np.int32)
    l_all = _c(np.concatenate([np.asarray(s["l"]) for s in species]), np.int32)
    dion = _c(np.concatenate([np.asarray(s["dion_ry"], np.float64).ravel() for s in species]), np.float64)
    n = lib().orc_d_blocks(len(asp), _ptr(asp), len(species), _ptr(nbeta), _ptr(l_all), _ptr(dion), None, None)
    sizes = np.zeros(len(asp), np.int32)
    out = np.zeros(max(n, 1))
    lib().orc_d_blocks(len(asp), _ptr(asp), len(species), _ptr(nbeta), _ptr(l_all), _ptr(dion), _ptr(sizes),
                       _ptr(out))
    return sizes, out[:n]


def zhegv(H, S):
    H = np.asfortranarray(H, dtype=np.complex128)
    S = np.asfortranarray(S, dtype=np.complex128)
    n = H.shape[0]
    w = np.zeros(n)
    X = np.zeros((n, n), np.complex128, order="F")
    info = lib().orc_zhegv(n, _ptr(H), _ptr(S), _ptr(w), _ptr(X))
    return info, w, X


def davidson(dims, fftmap, ekin, V, beta, block_sizes, dblocks, nb, nocc, psi0=None, tol=1e-8,
             max_iter=20, factor=2, minus_map=None, stored_eigenvalues=None):
    """Returns (eigenvalues, psi (npw x nb), report dict).  stored_eigenvalues: the k-point
    state's stored eigenvalues, the previous eigenvalues of the first convergence test when any
    is nonzero (davidson.hpp:473-493)."""
    dims, fm, ek, Vr, b, nproj, bs, db = _ham_args(dims, fftmap, ekin, V, beta, block_sizes, dblocks)
    npw = len(ek)
    if psi0 is None:
        psi = np.zeros((npw, nb), np.complex128, order="F")
        seed = 0
    else:
        psi = np.asfortranarray(psi0, dtype=np.complex128).copy(order="F")
        seed = 1
    eig = np.zeros(nb)
    if stored_eigenvalues is not None:
        eig[:] = np.asarray(stored_eigenvalues, dtype=np.float64).reshape(nb)
        seed |= 2
    rep = np.zeros(3, np.int32)
    mm = None if minus_map is None else _c(minus_map, np.int32)
    lib().orc_davidson2(_ptr(dims), npw, _ptr(fm), None if mm is None else _ptr(mm), _ptr(ek), _ptr(Vr), nproj,
                        _ptr(b), len(bs), _ptr(bs), _ptr(db), nb, nocc, _ptr(psi), seed, _ptr(eig), tol, max_iter,
                        factor, _ptr(rep))
    st = np.zeros(8)
    lib().orc_davidson_stage_ms(_ptr(st))
    stages = dict(zip(DAVIDSON_STAGES, (float(x) for x in st)))
    return eig, psi, {"iterations": int(rep[0]), "unconverged": int(rep[1]), "h_applications": int(rep[2]),
                      "stage_ms": stages}


# stage names of DavidsonTimingAccumulator (counters.hpp:36-44), in the order orc_davidson_stage_ms fills
DAVIDSON_STAGES = ("seed", "ritz_vectors", "residuals", "orthogonalize", "hamiltonian_application",
                   "subspace_update", "collapse", "diagonalize")


def set_threads(n: int) -> None:
    """OpenMP threads of the oracle's parallel loops (bands of H|psi>, GEMM columns)."""
    lib().orc_set_threads(int(n))


def get_threads() -> int:
    return int(lib().orc_get_threads())


# ---- the non-self-consistent k-loop over a band path
# (iterator/kohn_sham/planewave/pseudopotential/init.hpp:1036-1106, 1599-1815)

BACKWARD_SWEEP_DEPTH = 5  # pseudopotential/init.hpp:1767


def at_band_path_segment_boundary(k: int, segments: int, points_per_segment: int) -> bool:
    """is_at_band_path_segment_boundary (pseudopotential/init.hpp:1697-1715): k is a multiple of
    the points per segment below the last segment's end; never without a band path."""
    if segments <= 0 or points_per_segment <= 0:
        return False
    if k >= segments * points_per_segment:
        return False
    return k % points_per_segment == 0


def nscf_schedule(nk: int, segments: int = 0, points_per_segment: int = 0):
    """The solves of the k-loop in order, as (target k-point, seed k-point or None).  None: the
    k-point starts from its own initial state; a seed s: s's wavefunctions (truncated or
    zero-padded to the target's plane-wave count) and eigenvalues are copied in first.

    Forward pass (init.hpp:1056-1096): k = 0 .. nk-1, each seeded by k-1 unless k-1 is a segment
    boundary.  Backward fix (init.hpp:1735-1800), when nk > 2: for each edge (k-point 0, and the
    first k-point of every segment after the first), the k-points edge+5 .. edge (clipped to the
    path) are re-solved in descending order, the first from the k-point after them (or the last
    one), each next one from the one just solved."""
    steps = []
    for k in range(nk):
        seeded = k > 0 and not at_band_path_segment_boundary(k - 1, segments, points_per_segment)
        steps.append((k, k - 1 if seeded else None))
    if nk <= 2:
        return steps
    edges = [0]
    if segments > 0 and points_per_segment > 0:
        for s in range(1, segments):
            if s * points_per_segment < nk:
                edges.append(s * points_per_segment)
    for e in edges:
        end = min(e + BACKWARD_SWEEP_DEPTH, nk - 1)
        seed = min(end + 1, nk - 1)
        for t in range(end, e - 1, -1):
            steps.append((t, seed))
            seed = t
    return steps


def propagate(psi_src: np.ndarray, npw_target: int) -> np.ndarray:
    """propagate_wavefunctions_to_next_kpoint / _backward (init.hpp:1599-1688): the leading
    min(npw) coefficients of every band, zeros after them."""
    nb = psi_src.shape[1]
    out = np.zeros((npw_target, nb), np.complex128, order="F")
    n = min(psi_src.shape[0], npw_target)
    out[:n] = psi_src[:n]
    return out


def nscf_band_path(solve, initial, npw, nb, segments=0, points_per_segment=0):
    """Runs the k-loop over len(npw) k-points.  solve(k, psi0, stored_eigenvalues) -> (eig, psi,
    report); initial(k) -> the k-point's starting block (npw[k] x nb).  Stored eigenvalues start
    at zero.  Returns (eigenvalues nk x nb, reports in solve order, the schedule)."""
    nk = len(npw)
    psi = [None] * nk
    eig = np.zeros((nk, nb))
    reports = []
    steps = nscf_schedule(nk, segments, points_per_segment)
    for t, s in steps:
        if s is None:
            start = psi[t] if psi[t] is not None else initial(t)
        else:
            start = propagate(psi[s], npw[t])
            eig[t] = eig[s]
        e, p, rep = solve(t, start, eig[t].copy())
        eig[t] = e
        psi[t] = p
        reports.append(rep)
    return eig, reports, steps
