"""Host-side mirror of Q3's plane-wave Hamiltonian and Davidson call surface.

Same names, argument meaning and error behaviour as the reference's interfaces, over the
C ABI of libq3x.so:

* ``Crystal``                      -- crystal.hpp:156-177 (lattice rows in alat units)
* ``PlanewaveBasis``               -- basis/planewave/init.hpp:60-271 (generate, compute_gk)
* ``KPoint``                       -- the per-k list of PlaneWaveSelector::select
* ``PlanewaveHamiltonianApplicator`` -- hamiltonian_applicator/init.hpp (update_potential,
  prepare_kpoint, apply_batch; apply before prepare_kpoint raises like serial.hpp:487-492)
* ``PlanewaveDavidsonDiagonalizer`` -- diagonalizer/planewave/init.hpp:65-176 (set_tolerance,
  set_maximum_iterations, set_basis_expansion_factor, solve, number_of_unconverged_bands)

Arrays: complex band blocks are numpy ``complex128`` of shape (npw, nvec) in Fortran order
(column per band, ld = npw), exactly the reference's layout.  Energies in Ry.
"""
from __future__ import annotations

import atexit
import ctypes
import weakref
from dataclasses import dataclass, field

import numpy as np

from . import _lib
from ._lib import check, lib

FP64 = 0
MIXED = 1


def _ptr(a):
    return a.ctypes.data_as(ctypes.c_void_p)


# Device handles still open when the interpreter exits are released by an exit handler, which
# runs before the HIP runtime tears itself down, in a fixed order: servers and batches first
# (their worker threads use contexts of their own), then applicators and device arrays, then
# contexts.  Left to garbage collection the order is arbitrary.
_open = (weakref.WeakSet(), weakref.WeakSet(), weakref.WeakSet())


def _track(obj, rank: int):
    _open[rank].add(obj)
    return obj


@atexit.register
def _release_open_handles():
    for group in _open:
        for obj in list(group):
            try:
                obj.close()
            except Exception:  # noqa: BLE001 - best effort at exit
                pass


@dataclass
class Crystal:
    """Lattice rows a1, a2, a3 in alat units, alat in bohr; atoms in crystal coordinates."""
    lattice: np.ndarray
    alat: float
    positions: np.ndarray = field(default_factory=lambda: np.zeros((0, 3)))
    species: np.ndarray = field(default_factory=lambda: np.zeros(0, np.int32))

    def cartesian_positions_bohr(self):
        return (np.asarray(self.positions) @ np.asarray(self.lattice)) * self.alat

    @property
    def number_of_atoms(self):
        return len(self.positions)


class Context:
    """One device and one HIP stream (default priority)."""

    def __init__(self, device: int = 0):
        h = ctypes.c_void_p()
        check(lib().q3x_context_create(device, ctypes.byref(h)))
        self.h = h
        _track(self, 2)

    def synchronize(self):
        check(lib().q3x_context_synchronize(self.h))

    def close(self):
        if getattr(self, "h", None):
            lib().q3x_context_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()


class DeviceArray:
    """A device allocation of the context (complex128 band blocks resident in HBM)."""

    def __init__(self, context: Context, nbytes: int):
        self.context = context
        self.nbytes = int(nbytes)
        p = ctypes.c_void_p()
        check(lib().q3x_device_alloc(context.h, self.nbytes, ctypes.byref(p)))
        self.ptr = p.value
        _track(self, 1)

    @classmethod
    def from_host(cls, context: Context, a: np.ndarray):
        a = np.asfortranarray(a)
        d = cls(context, a.nbytes)
        d.upload(a)
        return d

    def upload(self, a: np.ndarray):
        a = np.asfortranarray(a)
        assert a.nbytes <= self.nbytes
        check(lib().q3x_copy_to_device(self.context.h, ctypes.c_void_p(self.ptr), _ptr(a), a.nbytes))

    def download(self, shape, dtype=np.complex128) -> np.ndarray:
        out = np.zeros(shape, dtype, order="F")
        assert out.nbytes <= self.nbytes
        check(lib().q3x_copy_to_host(self.context.h, _ptr(out), ctypes.c_void_p(self.ptr), out.nbytes))
        return out

    def free(self):
        if getattr(self, "ptr", None) and getattr(self.context, "h", None):
            lib().q3x_device_free(self.context.h, ctypes.c_void_p(self.ptr))
        self.ptr = None

    def close(self):
        self.free()

    def __del__(self):
        self.free()


def zgemm(context: Context, transa: str, m: int, n: int, k: int, alpha: float, A: DeviceArray, lda: int,
          B: DeviceArray, ldb: int, beta: float, C: DeviceArray, ldc: int, precision: int = FP64):
    """C = alpha op(A) B + beta C on device buffers (MFMA complex GEMM)."""
    check(lib().q3x_zgemm(context.h, transa.encode(), m, n, k, alpha, ctypes.c_void_p(A.ptr), lda,
                          ctypes.c_void_p(B.ptr), ldb, beta, ctypes.c_void_p(C.ptr), ldc, precision))


def hegv_device(context: Context, n: int, H: DeviceArray, ldh: int, S: DeviceArray, lds: int, nvec: int,
                X: DeviceArray, ldx: int) -> np.ndarray:
    """H x = w S x on device matrices (upper triangles read); returns the nvec lowest eigenvalues
    (ascending), the S-orthonormal eigenvectors land in X (n x nvec, ld ldx)."""
    w = np.zeros(max(nvec, 1))
    check(lib().q3x_hegv_device(context.h, n, ctypes.c_void_p(H.ptr), ldh, ctypes.c_void_p(S.ptr), lds, nvec,
                                _ptr(w), ctypes.c_void_p(X.ptr), ldx))
    return w[:nvec]


class PlanewaveBasis:
    def __init__(self, crystal: Crystal, ecutwfc_ry: float, ecutrho_ry: float = 0.0,
                 use_exact_wavefunction_grid: bool = True, fft_factors=(1, 1, 1), gamma_only: bool = False):
        """gamma_only: the Gamma trick (half G sphere, real bands at k = 0), as the reference uses for a
        1x1x1 unshifted k grid (iterator/kohn_sham/planewave/init.hpp:585-610)."""
        self.crystal = crystal
        self.gamma_only = bool(gamma_only)
        lat = np.ascontiguousarray(crystal.lattice, dtype=np.float64).ravel()
        fac = np.ascontiguousarray(fft_factors, dtype=np.int32)
        h = ctypes.c_void_p()
        check(lib().q3x_basis_create_gamma(_ptr(lat), float(crystal.alat), float(ecutwfc_ry), float(ecutrho_ry),
                                           int(use_exact_wavefunction_grid), _ptr(fac), int(gamma_only),
                                           ctypes.byref(h)))
        self.h = h
        rho = np.zeros(3, np.int32)
        wave = np.zeros(3, np.int32)
        ng = np.zeros(1, np.int32)
        vol = np.zeros(1)
        tpa2 = np.zeros(1)
        check(lib().q3x_basis_info(h, _ptr(rho), _ptr(wave), _ptr(ng), _ptr(vol), _ptr(tpa2)))
        self.rho_dims = tuple(int(x) for x in rho)
        self.wave_dims = tuple(int(x) for x in wave)
        self.number_of_g_vectors = int(ng[0])
        self.cell_volume = float(vol[0])
        self.two_pi_over_alat_squared = float(tpa2[0])
        self.ecutwfc = ecutwfc_ry
        self.ecutrho = ecutrho_ry if ecutrho_ry > 0 else 4.0 * ecutwfc_ry

    def g_vectors(self):
        n = self.number_of_g_vectors
        m = np.zeros((n, 3), np.int32)
        g = np.zeros((n, 3))
        g2 = np.zeros(n)
        check(lib().q3x_basis_gvectors(self.h, _ptr(m), _ptr(g), _ptr(g2)))
        return m, g, g2

    def compute_gk(self, k_crystal):
        return KPoint(self, k_crystal)

    def __del__(self):
        if getattr(self, "h", None):
            lib().q3x_basis_destroy(self.h)
            self.h = None


class KPoint:
    def __init__(self, basis: PlanewaveBasis, k_crystal):
        self.basis = basis
        self.k = np.ascontiguousarray(k_crystal, dtype=np.float64)
        h = ctypes.c_void_p()
        npw = np.zeros(1, np.int32)
        check(lib().q3x_kpoint_create(basis.h, _ptr(self.k), ctypes.byref(h), _ptr(npw)))
        self.h = h
        self.npw = int(npw[0])
        self.g_vector_indices = np.zeros(self.npw, np.int32)
        self.kinetic_energy_squared = np.zeros(self.npw)
        self.fft_map = np.zeros(self.npw, np.int32)
        check(lib().q3x_kpoint_maps(h, _ptr(self.g_vector_indices), _ptr(self.kinetic_energy_squared),
                                    _ptr(self.fft_map)))
        self.minus_fft_map = None
        if basis.gamma_only:
            self.minus_fft_map = np.zeros(self.npw, np.int32)
            check(lib().q3x_kpoint_minus_map(h, _ptr(self.minus_fft_map)))

    def random_psi(self, nbands: int, seed: int) -> np.ndarray:
        """Starting block (npw x nbands, Fortran order): entries uniform in the unit square damped
        by 1 / (1 + |k+G|^2), the generator the band-structure server seeds k-point i with (seed i)."""
        out = np.zeros((self.npw, nbands), np.complex128, order="F")
        check(lib().q3x_kpoint_random_psi(self.h, nbands, seed, _ptr(out)))
        return out

    def kb_projectors(self, crystal: Crystal, species):
        """Projector set <k+G|beta> (npw x nproj, Fortran order).  species: list of dicts with
        r, rab, l (list), cutoff (list), beta_r (nbeta x nmesh)."""
        pos = np.ascontiguousarray(crystal.positions, dtype=np.float64)
        asp = np.ascontiguousarray(crystal.species, dtype=np.int32)
        nmesh = np.ascontiguousarray([len(s["r"]) for s in species], dtype=np.int32)
        r_all = np.ascontiguousarray(np.concatenate([s["r"] for s in species]), dtype=np.float64)
        rab_all = np.ascontiguousarray(np.concatenate([s["rab"] for s in species]), dtype=np.float64)
        nbeta = np.ascontiguousarray([len(s["l"]) for s in species], dtype=np.int32)
        l_all = np.ascontiguousarray(np.concatenate([s["l"] for s in species]), dtype=np.int32)
        cut_all = np.ascontiguousarray(np.concatenate([s["cutoff"] for s in species]), dtype=np.int32)
        br_all = np.ascontiguousarray(np.concatenate([np.asarray(s["beta_r"]).ravel() for s in species]),
                                      dtype=np.float64)
        nproj = np.zeros(1, np.int32)
        args = (self.basis.h, self.h, len(asp), _ptr(pos), _ptr(asp), len(species), _ptr(nmesh), _ptr(r_all),
                _ptr(rab_all), _ptr(nbeta), _ptr(l_all), _ptr(cut_all), _ptr(br_all))
        check(lib().q3x_kb_projectors(*args, None, _ptr(nproj)))
        out = np.zeros((int(nproj[0]), self.npw), np.complex128)
        check(lib().q3x_kb_projectors(*args, _ptr(out), _ptr(nproj)))
        return np.asfortranarray(out.T)

    def __del__(self):
        if getattr(self, "h", None):
            lib().q3x_kpoint_destroy(self.h)
            self.h = None


class PlanewaveHamiltonianApplicator:
    """H|psi> on the GPU for one prepared k-point (serial.hpp semantics)."""

    def __init__(self, context: Context, basis: PlanewaveBasis):
        self.context = context
        self.basis = basis
        h = ctypes.c_void_p()
        check(lib().q3x_hamiltonian_create(context.h, basis.h, ctypes.byref(h)))
        self.h = h
        self.kpoint = None
        self.nproj = 0
        _track(self, 1)

    def update_potential(self, v_r: np.ndarray):
        v = np.ascontiguousarray(v_r, dtype=np.float64).ravel()
        check(lib().q3x_hamiltonian_set_potential(self.h, _ptr(v), v.size))

    def prepare_kpoint(self, kpoint: KPoint):
        check(lib().q3x_hamiltonian_prepare_kpoint(self.h, kpoint.h))
        self.kpoint = kpoint
        self.nproj = 0

    def build_projectors(self, crystal: Crystal, species, block_sizes, d_blocks):
        """The projector set of the prepared k-point built on the device from the species (the
        arrays kb_projectors takes) and the D blocks; no host projector array exists."""
        pos = np.ascontiguousarray(crystal.positions, dtype=np.float64)
        asp = np.ascontiguousarray(crystal.species, dtype=np.int32)
        nmesh = np.ascontiguousarray([len(s["r"]) for s in species], dtype=np.int32)
        r_all = np.ascontiguousarray(np.concatenate([s["r"] for s in species]), dtype=np.float64)
        rab_all = np.ascontiguousarray(np.concatenate([s["rab"] for s in species]), dtype=np.float64)
        nbeta = np.ascontiguousarray([len(s["l"]) for s in species], dtype=np.int32)
        l_all = np.ascontiguousarray(np.concatenate([s["l"] for s in species]), dtype=np.int32)
        cut_all = np.ascontiguousarray(np.concatenate([s["cutoff"] for s in species]), dtype=np.int32)
        br_all = np.ascontiguousarray(np.concatenate([np.asarray(s["beta_r"]).ravel() for s in species]),
                                      dtype=np.float64)
        bs = np.ascontiguousarray(block_sizes, dtype=np.int32)
        db = np.ascontiguousarray(d_blocks, dtype=np.float64)
        check(lib().q3x_hamiltonian_build_projectors(self.h, len(asp), _ptr(pos), _ptr(asp), len(species),
                                                     _ptr(nmesh), _ptr(r_all), _ptr(rab_all), _ptr(nbeta),
                                                     _ptr(l_all), _ptr(cut_all), _ptr(br_all), _ptr(bs), _ptr(db)))
        self.nproj = int(bs.sum())

    def random_psi_device(self, nbands: int, seed: int, d_psi: int):
        """The starting block of KPoint.random_psi(nbands, seed) written on the device."""
        check(lib().q3x_hamiltonian_random_psi(self.h, nbands, seed, ctypes.c_void_p(d_psi)))

    def set_projectors(self, beta: np.ndarray, block_sizes, d_blocks):
        b = np.asfortranarray(beta, dtype=np.complex128)
        bs = np.ascontiguousarray(block_sizes, dtype=np.int32)
        db = np.ascontiguousarray(d_blocks, dtype=np.float64)
        nproj = b.shape[1] if b.ndim == 2 else 0
        check(lib().q3x_hamiltonian_set_projectors(self.h, _ptr(b), nproj, len(bs), _ptr(bs), _ptr(db)))
        self.nproj = nproj

    def apply_batch(self, psi: np.ndarray, precision: int = FP64) -> np.ndarray:
        psi = np.asfortranarray(psi, dtype=np.complex128)
        if psi.ndim == 1:
            psi = psi[:, None]
        out = np.zeros_like(psi, order="F")
        check(lib().q3x_hamiltonian_apply(self.h, _ptr(psi), _ptr(out), psi.shape[1], precision))
        return out

    def apply_device(self, d_psi: int, d_hpsi: int, nvec: int, precision: int = FP64):
        """psi / hpsi already resident in HBM (device pointers, complex128, ld = npw)."""
        check(lib().q3x_hamiltonian_apply_device(self.h, ctypes.c_void_p(d_psi), ctypes.c_void_p(d_hpsi), nvec,
                                                 precision))

    def accumulate_density_device(self, d_psi: int, nbands: int, weights, d_rho: int):
        """rho(r) += sum_b weights[b] |psi_b(r)|^2 on the wavefunction box (device pointers: psi
        npw x nbands complex128, rho n1 n2 n3 doubles, C order), stream-ordered
        (accumulator.hpp:159-280)."""
        w = np.ascontiguousarray(weights, dtype=np.float64).ravel()
        if w.size != nbands:
            raise ValueError("one weight per band")
        check(lib().q3x_hamiltonian_accumulate_density(self.h, ctypes.c_void_p(d_psi), nbands, _ptr(w),
                                                       ctypes.c_void_p(d_rho)))

    def density(self, psi: np.ndarray, weights) -> np.ndarray:
        """sum_b weights[b] |psi_b(r)|^2 of host bands (npw x nb) on the wavefunction box, as an
        (n1, n2, n3) array."""
        psi = np.asfortranarray(psi, dtype=np.complex128)
        if psi.ndim == 1:
            psi = psi[:, None]
        dims = tuple(self.basis.wave_dims)
        d_psi = DeviceArray.from_host(self.context, psi)
        d_rho = DeviceArray.from_host(self.context, np.zeros(int(np.prod(dims))))
        try:
            self.accumulate_density_device(d_psi.ptr, psi.shape[1], weights, d_rho.ptr)
            self.context.synchronize()
            return d_rho.download((int(np.prod(dims)),), np.float64).reshape(dims)
        finally:
            d_psi.free()
            d_rho.free()

    def apply_noalloc(self, d_psi: int, d_hpsi: int, nvec: int, precision: int = FP64):
        """apply_device with every scratch buffer in the library's fixed per-GPU array: no device
        allocation (for callers that cannot allocate).  Calls overlapping on one GPU race."""
        check(lib().q3x_hamiltonian_apply_noalloc(self.h, ctypes.c_void_p(d_psi), ctypes.c_void_p(d_hpsi), nvec,
                                                  precision))

    def close(self):
        if getattr(self, "h", None):
            lib().q3x_hamiltonian_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()


class PlanewaveDavidsonDiagonalizer:
    """Block Davidson on the GPU (davidson.hpp:324-841 control flow)."""

    def __init__(self, applicator: PlanewaveHamiltonianApplicator):
        self.applicator = applicator
        self.options = _lib.DavidsonOptions()
        lib().q3x_davidson_default_options(ctypes.byref(self.options))
        self.report = _lib.DavidsonReport()

    def set_tolerance(self, tol_ry: float):
        self.options.tolerance_ry = tol_ry

    def set_maximum_iterations(self, n: int):
        self.options.max_iterations = n

    def set_basis_expansion_factor(self, f: int):
        self.options.basis_factor = f

    def number_of_unconverged_bands(self) -> int:
        return self.report.unconverged

    def solve(self, number_of_bands: int, number_of_occupied: int, psi0: np.ndarray | None = None,
              precision: int = FP64, final_fp64_sweep: bool = True, stored_eigenvalues=None):
        """Returns (eigenvalues [Ry], psi (npw x nb)).  stored_eigenvalues: the k-point state's
        stored eigenvalues (nb, Ry), the previous eigenvalues of the first convergence test when
        any is nonzero (davidson.hpp:473-493)."""
        kp = self.applicator.kpoint
        if kp is None:
            raise _lib.StateError(3, "prepare_kpoint() must be called before solve")
        nb = number_of_bands
        self.options.n_occupied = number_of_occupied
        self.options.precision = precision
        self.options.final_fp64_sweep = int(final_fp64_sweep)
        if psi0 is None:
            psi = np.zeros((kp.npw, nb), np.complex128, order="F")
            self.options.seed_from_psi = 0
        else:
            psi = np.asfortranarray(psi0, dtype=np.complex128).copy(order="F")
            self.options.seed_from_psi = 1
        eig = np.zeros(nb)
        self.options.stored_eigenvalues = int(stored_eigenvalues is not None)
        if stored_eigenvalues is not None:
            eig[:] = np.asarray(stored_eigenvalues, dtype=np.float64).reshape(nb)
        self.report = _lib.DavidsonReport()
        check(lib().q3x_davidson_solve(self.applicator.h, nb, ctypes.byref(self.options), _ptr(psi), _ptr(eig),
                                       ctypes.byref(self.report)))
        return eig, psi

    def report_dict(self):
        r = self.report
        names = ["seed", "orthogonalize", "hamiltonian_application", "subspace_update", "diagonalize",
                 "collapse", "residual", "total"]
        return {"iterations": r.iterations, "unconverged": r.unconverged, "h_applications": r.h_applications,
                "fp64_refinements": r.fp64_refinements, "max_residual_norm": r.max_residual_norm,
                "eig_barrier_redos": r.eig_barrier_redos, "mgs_fallbacks": r.mgs_fallbacks,
                "recovery_collapses": r.recovery_collapses, "previous_eigenvalue_keeps": r.previous_eigenvalue_keeps,
                "stage_ms": {n: r.stage_ms[i] for i, n in enumerate(names)}}


class KPointBatch:
    """Independent k-points solved together on one GPU (replaces the SCF k-loop of
    iterator/kohn_sham/planewave/pseudopotential/init.hpp:1056-1082 for independent k-points):
    k-point i runs on worker i % workers, each worker with its own stream and host thread."""

    def __init__(self, workers: int = 8, device: int = 0):
        h = ctypes.c_void_p()
        check(lib().q3x_batch_create(device, workers, ctypes.byref(h)))
        self.h = h
        _track(self, 0)
        self.workers = workers
        self.npw = []
        self._keep = []  # the batch's Hamiltonians refer to these bases and k-point lists
        self.options = _lib.DavidsonOptions()
        lib().q3x_davidson_default_options(ctypes.byref(self.options))
        self.reports = []

    def add(self, basis: PlanewaveBasis, kpoint: KPoint, v_r: np.ndarray, beta: np.ndarray, block_sizes,
            d_blocks) -> int:
        v = np.ascontiguousarray(v_r, dtype=np.float64).ravel()
        b = np.asfortranarray(beta, dtype=np.complex128)
        bs = np.ascontiguousarray(block_sizes, dtype=np.int32)
        db = np.ascontiguousarray(d_blocks, dtype=np.float64)
        idx = ctypes.c_int()
        check(lib().q3x_batch_add(self.h, basis.h, kpoint.h, _ptr(v), v.size, _ptr(b), b.shape[1], len(bs), _ptr(bs),
                                  _ptr(db), ctypes.byref(idx)))
        self.npw.append(kpoint.npw)
        self._keep.append((basis, kpoint))
        return idx.value

    def solve(self, number_of_bands: int, number_of_occupied: int, tolerance_ry: float, max_iterations: int = 20,
              precision: int = FP64, psi0=None, final_fp64_sweep: bool = True):
        """Returns (eigenvalues nk x nb [Ry], list of psi_k (npw_k x nb))."""
        nb = number_of_bands
        o = self.options
        o.tolerance_ry = tolerance_ry
        o.max_iterations = max_iterations
        o.n_occupied = number_of_occupied
        o.precision = precision
        o.final_fp64_sweep = int(final_fp64_sweep)
        o.seed_from_psi = int(psi0 is not None)
        if psi0 is None:
            psis = [np.zeros((n, nb), np.complex128, order="F") for n in self.npw]
        else:
            psis = [np.asfortranarray(p, dtype=np.complex128).copy(order="F") for p in psi0]
        ptrs = (ctypes.c_void_p * len(psis))(*[p.ctypes.data for p in psis])
        eig = np.zeros((len(psis), nb))
        reps = (_lib.DavidsonReport * len(psis))()
        check(lib().q3x_batch_solve(self.h, nb, ctypes.byref(o), ptrs, _ptr(eig), reps))
        self.reports = list(reps)
        return eig, psis

    def close(self):
        if getattr(self, "h", None):
            lib().q3x_batch_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()


def _species_arrays(species):
    nmesh = np.ascontiguousarray([len(s["r"]) for s in species], dtype=np.int32)
    return {
        "nmesh": nmesh,
        "r_all": np.ascontiguousarray(np.concatenate([s["r"] for s in species]), dtype=np.float64),
        "rab_all": np.ascontiguousarray(np.concatenate([s["rab"] for s in species]), dtype=np.float64),
        "nbeta": np.ascontiguousarray([len(s["l"]) for s in species], dtype=np.int32),
        "l_all": np.ascontiguousarray(np.concatenate([s["l"] for s in species]), dtype=np.int32),
        "cutoff_all": np.ascontiguousarray(np.concatenate([s["cutoff"] for s in species]), dtype=np.int32),
        "beta_r_all": np.ascontiguousarray(np.concatenate([np.asarray(s["beta_r"]).ravel() for s in species]),
                                           dtype=np.float64),
    }


@dataclass
class BandRequestSpec:
    """One band-structure request: a cell (basis), its V(r), the k-points and the band count.
    kpath_propagation: 0 independent k-points; 1 the reference's non-self-consistent k-loop
    (forward propagation along the path, segment boundaries, the backward fix), with
    band_path_segments segments of band_path_points_per_segment points (0: no band path)."""
    basis: PlanewaveBasis
    v_r: np.ndarray
    kpoints: np.ndarray
    species: list
    block_sizes: np.ndarray
    d_blocks: np.ndarray
    nbands: int
    noccupied: int
    tolerance_ry: float = 2e-6
    max_iterations: int = 60
    precision: int = FP64
    kpath_propagation: int = 0
    band_path_segments: int = 0
    band_path_points_per_segment: int = 0


def _band_request(spec: BandRequestSpec):
    """The q3x_band_request of a spec and the arrays it points into (keep them alive)."""
    c = spec.basis.crystal
    arrays = {
        "v": np.ascontiguousarray(spec.v_r, dtype=np.float64).ravel(),
        "k": np.ascontiguousarray(spec.kpoints, dtype=np.float64).reshape(-1, 3),
        "pos": np.ascontiguousarray(c.positions, dtype=np.float64),
        "sp": np.ascontiguousarray(c.species, dtype=np.int32),
        "bs": np.ascontiguousarray(spec.block_sizes, dtype=np.int32),
        "db": np.ascontiguousarray(spec.d_blocks, dtype=np.float64),
        **_species_arrays(spec.species),
    }
    r = _lib.BandRequest()
    r.basis = spec.basis.h
    r.v_r = arrays["v"].ctypes.data
    r.nv = arrays["v"].size
    r.nk = arrays["k"].shape[0]
    r.kpoints = arrays["k"].ctypes.data
    r.natoms = len(arrays["sp"])
    r.atom_pos = arrays["pos"].ctypes.data
    r.atom_species = arrays["sp"].ctypes.data
    r.nspecies = len(spec.species)
    for name in ("nmesh", "r_all", "rab_all", "nbeta", "l_all", "cutoff_all", "beta_r_all"):
        setattr(r, name, arrays[name].ctypes.data)
    r.block_sizes = arrays["bs"].ctypes.data
    r.d_blocks = arrays["db"].ctypes.data
    r.nbands = spec.nbands
    r.noccupied = spec.noccupied
    lib().q3x_davidson_default_options(ctypes.byref(r.options))
    r.options.tolerance_ry = spec.tolerance_ry
    r.options.max_iterations = spec.max_iterations
    r.options.precision = spec.precision
    r.kpath_propagation = spec.kpath_propagation
    r.band_path_segments = spec.band_path_segments
    r.band_path_points_per_segment = spec.band_path_points_per_segment
    return r, arrays


class BandStructureServer:
    """Continuous-batching server of band-structure requests (q3x_server_*): requests are
    admitted in arrival order while their estimated HBM footprint fits the budget, and the
    k-points of all admitted requests share the worker streams."""

    def __init__(self, workers: int = 8, hbm_budget_bytes: float = 0.0, device: int = 0):
        h = ctypes.c_void_p()
        check(lib().q3x_server_create(device, workers, float(hbm_budget_bytes), ctypes.byref(h)))
        self.h = h
        _track(self, 0)
        self._live = {}  # ticket -> (spec, arrays) kept alive until the result is collected

    def _request(self, spec: BandRequestSpec):
        return _band_request(spec)

    def estimate(self, spec: BandRequestSpec) -> float:
        r, _arrays = self._request(spec)
        out = ctypes.c_double()
        check(lib().q3x_server_estimate(self.h, ctypes.byref(r), ctypes.byref(out)))
        return out.value

    def submit(self, spec: BandRequestSpec) -> int:
        r, arrays = self._request(spec)
        t = ctypes.c_long()
        check(lib().q3x_server_submit(self.h, ctypes.byref(r), ctypes.byref(t)))
        self._live[t.value] = (spec, arrays)
        return t.value

    def wait(self, ticket: int):
        """Returns (eigenvalues nk x nb [Ry], stats dict) and frees the request's device state;
        raises if the request failed or was cancelled."""
        status, eig, stats = self.wait_status(ticket)
        check(status)
        return eig, stats

    def wait_status(self, ticket: int):
        """Like wait() but returns (status, eigenvalues, stats) instead of raising; the stats of
        a cancelled request count the k-points it solved."""
        spec, arrays = self._live[ticket]
        eig = np.zeros((arrays["k"].shape[0], spec.nbands))
        st = _lib.RequestStats()
        try:
            status = lib().q3x_server_wait(self.h, ticket, _ptr(eig), ctypes.byref(st))
        finally:
            del self._live[ticket]
        return status, eig, {"submit_s": st.submit_s, "admit_s": st.admit_s, "done_s": st.done_s,
                             "footprint_bytes": st.footprint_bytes, "iterations": st.iterations,
                             "setup_s": st.setup_s, "solve_s": st.solve_s, "eig_barrier_redos": st.eig_barrier_redos,
                             "kpoints_done": st.kpoints_done, "solves": st.solves}

    def in_flight_bytes(self) -> float:
        return lib().q3x_server_in_flight_bytes(self.h)

    def cancel_pending(self) -> int:
        """Drops every k-point not yet started; affected requests then fail with StateError in
        wait().  Returns the number of requests affected."""
        n = ctypes.c_int(0)
        check(lib().q3x_server_cancel_pending(self.h, ctypes.byref(n)))
        return n.value

    def close(self):
        if getattr(self, "h", None):
            lib().q3x_server_destroy(self.h)
            self.h = None
            self._live.clear()

    def __del__(self):
        self.close()


def memory_estimate(**inputs) -> dict:
    """PlanewaveMemoryEstimate::compute (iterator/kohn_sham/planewave/internal/memory_estimate.hpp:
    75-215) in the library: keyword arguments are the reference's input fields (defaults as
    PlanewaveMemoryEstimateInputs{}); returns the terms and totals."""
    x = _lib.MemoryEstimateInputs()
    lib().q3x_memory_estimate_default_inputs(ctypes.byref(x))
    for k, v in inputs.items():
        if not hasattr(x, k):
            raise TypeError(f"unknown input {k!r}")
        setattr(x, k, int(v))
    e = _lib.MemoryEstimate()
    check(lib().q3x_memory_estimate_compute(ctypes.byref(x), ctypes.byref(e)))
    return _estimate_dict(e)


def _terms(arr, n):
    return [(t.label.decode(), t.formula.decode(), int(t.bytes)) for t in arr[:n]]


def _estimate_dict(e) -> dict:
    return {"steady_state": _terms(e.steady_state, e.n_steady_state), "transient": _terms(e.transient, e.n_transient),
            "steady_state_bytes": int(e.steady_state_bytes), "solver_bytes": int(e.solver_bytes),
            "setup_peak_bytes": int(e.setup_peak_bytes), "peak_bytes": int(e.peak_bytes),
            "maximum_basis_size": int(e.maximum_basis_size)}


def band_request_footprint(spec: BandRequestSpec, concurrent_kpoints: int) -> dict:
    """The server's admission footprint of a request, itemised (q3x_band_request_footprint):
    {"bytes", "reference": the reference estimate of its dimensions, "device_terms"}."""
    r, _arrays = _band_request(spec)
    e = _lib.MemoryEstimate()
    terms = (_lib.MemoryTerm * 8)()
    n = ctypes.c_int(0)
    b = ctypes.c_double()
    check(lib().q3x_band_request_footprint(ctypes.byref(r), int(concurrent_kpoints), ctypes.byref(e), terms,
                                           ctypes.byref(n), ctypes.byref(b)))
    return {"bytes": b.value, "reference": _estimate_dict(e), "device_terms": _terms(terms, n.value)}
