"""ctypes binding of libq3x.so (the C ABI declared in include/q3x.h).

The library is loaded from the package directory only; if it is missing this raises
instead of falling back to anything else.
"""
from __future__ import annotations

import ctypes
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libq3x.so")

_i = ctypes.c_int
_l = ctypes.c_long
_ul = ctypes.c_ulong
_d = ctypes.c_double
_p = ctypes.c_void_p
_pp = ctypes.POINTER(ctypes.c_void_p)


class DavidsonOptions(ctypes.Structure):
    _fields_ = [("tolerance_ry", _d), ("max_iterations", _i), ("basis_factor", _i), ("n_occupied", _i),
                ("precision", _i), ("seed_from_psi", _i), ("final_fp64_sweep", _i),
                ("stored_eigenvalues", _i)]


class DavidsonReport(ctypes.Structure):
    _fields_ = [("iterations", _i), ("unconverged", _i), ("h_applications", _i), ("fp64_refinements", _i),
                ("max_residual_norm", _d), ("stage_ms", _d * 8), ("eig_barrier_redos", _i), ("mgs_fallbacks", _i),
                ("recovery_collapses", _i), ("previous_eigenvalue_keeps", _i)]


class BandRequest(ctypes.Structure):
    _fields_ = [("basis", _p), ("v_r", _p), ("nv", _l), ("nk", _i), ("kpoints", _p), ("natoms", _i),
                ("atom_pos", _p), ("atom_species", _p), ("nspecies", _i), ("nmesh", _p), ("r_all", _p),
                ("rab_all", _p), ("nbeta", _p), ("l_all", _p), ("cutoff_all", _p), ("beta_r_all", _p),
                ("block_sizes", _p), ("d_blocks", _p), ("nbands", _i), ("noccupied", _i),
                ("options", DavidsonOptions), ("kpath_propagation", _i), ("band_path_segments", _i),
                ("band_path_points_per_segment", _i)]


class MemoryEstimateInputs(ctypes.Structure):
    _fields_ = [("number_of_bands", _i), ("max_plane_waves_per_kpoint", _i), ("number_of_g_vectors", _i),
                ("number_of_projectors", _i), ("number_of_atomic_orbitals", _i), ("number_of_atoms", _i),
                ("number_of_spin_components", _i), ("local_number_of_kpoint_states", _i),
                ("mixing_history_depth", _i), ("fft_rho_grid_points", ctypes.c_longlong),
                ("fft_wave_grid_points", ctypes.c_longlong), ("davidson_basis_expansion_factor", _i),
                ("overlap_is_identity", _i), ("diagonalizer_is_davidson", _i)]


class MemoryTerm(ctypes.Structure):
    _fields_ = [("label", ctypes.c_char_p), ("formula", ctypes.c_char_p), ("bytes", ctypes.c_ulonglong)]


class MemoryEstimate(ctypes.Structure):
    _fields_ = [("steady_state", MemoryTerm * 16), ("n_steady_state", _i), ("transient", MemoryTerm * 4),
                ("n_transient", _i), ("steady_state_bytes", ctypes.c_ulonglong),
                ("solver_bytes", ctypes.c_ulonglong), ("setup_peak_bytes", ctypes.c_ulonglong),
                ("peak_bytes", ctypes.c_ulonglong), ("maximum_basis_size", _i)]


class RequestStats(ctypes.Structure):
    _fields_ = [("submit_s", _d), ("admit_s", _d), ("done_s", _d), ("footprint_bytes", _d), ("iterations", _i),
                ("setup_s", _d), ("solve_s", _d), ("eig_barrier_redos", _i), ("kpoints_done", _i),
                ("solves", _i)]


# name -> (restype, argtypes); every symbol declared in include/q3x.h
SIGNATURES = {
    "q3x_last_error": (ctypes.c_char_p, []),
    "q3x_version": (ctypes.c_char_p, []),
    "q3x_context_create": (_i, [_i, _pp]),
    "q3x_context_destroy": (_i, [_p]),
    "q3x_context_synchronize": (_i, [_p]),
    "q3x_basis_create": (_i, [_p, _d, _d, _d, _i, _p, _pp]),
    "q3x_basis_info": (_i, [_p, _p, _p, _p, _p, _p]),
    "q3x_basis_gvectors": (_i, [_p, _p, _p, _p]),
    "q3x_basis_destroy": (_i, [_p]),
    "q3x_kpoint_create": (_i, [_p, _p, _pp, _p]),
    "q3x_kpoint_maps": (_i, [_p, _p, _p, _p]),
    "q3x_kpoint_random_psi": (_i, [_p, _i, ctypes.c_ulonglong, _p]),
    "q3x_kpoint_destroy": (_i, [_p]),
    "q3x_kb_projectors": (_i, [_p, _p, _i, _p, _p, _i, _p, _p, _p, _p, _p, _p, _p, _p, _p]),
    "q3x_kb_d_blocks": (_i, [_i, _p, _i, _p, _p, _p, _p, _p, ctypes.POINTER(_l)]),
    "q3x_hamiltonian_create": (_i, [_p, _p, _pp]),
    "q3x_hamiltonian_destroy": (_i, [_p]),
    "q3x_hamiltonian_set_potential": (_i, [_p, _p, _l]),
    "q3x_hamiltonian_prepare_kpoint": (_i, [_p, _p]),
    "q3x_hamiltonian_set_projectors": (_i, [_p, _p, _i, _i, _p, _p]),
    "q3x_hamiltonian_build_projectors": (_i, [_p, _i, _p, _p, _i, _p, _p, _p, _p, _p, _p, _p, _p, _p]),
    "q3x_hamiltonian_random_psi": (_i, [_p, _i, ctypes.c_ulonglong, _p]),
    "q3x_hamiltonian_apply": (_i, [_p, _p, _p, _i, _i]),
    "q3x_hamiltonian_apply_device": (_i, [_p, _p, _p, _i, _i]),
    "q3x_hamiltonian_accumulate_density": (_i, [_p, _p, _i, _p, _p]),
    "q3x_hamiltonian_apply_noalloc": (_i, [_p, _p, _p, _i, _i]),
    "q3x_noalloc_scratch_bytes": (ctypes.c_ulong, []),
    "q3x_basis_create_gamma": (_i, [_p, _d, _d, _d, _i, _p, _i, _pp]),
    "q3x_basis_is_gamma_only": (_i, [_p, _p]),
    "q3x_kpoint_minus_map": (_i, [_p, _p]),
    "q3x_debug_force_barrier_giveups": (None, [_i]),
    "q3x_debug_fail_eigensolves": (None, [ctypes.c_ulonglong]),
    "q3x_eig_barrier_redos_total": (_l, []),
    "q3x_zgemm": (_i, [_p, ctypes.c_char, _i, _i, _i, _d, _p, _l, _p, _l, _d, _p, _l, _i]),
    "q3x_device_alloc": (_i, [_p, _ul, _pp]),
    "q3x_device_free": (_i, [_p, _p]),
    "q3x_device_copy": (_i, [_p, _p, _p, _ul]),
    "q3x_copy_to_device": (_i, [_p, _p, _p, _ul]),
    "q3x_copy_to_host": (_i, [_p, _p, _p, _ul]),
    "q3x_profile_enable": (_i, [_p, _i]),
    "q3x_profile_read": (_i, [_p, _i, _p, _p, _i]),
    "q3x_profile_kernel_name": (ctypes.c_char_p, [_i]),
    "q3x_hegv": (_i, [_i, _p, _p, _p, _p, _i]),
    "q3x_hegv_device": (_i, [_p, _i, _p, _l, _p, _l, _i, _p, _p, _l]),
    "q3x_davidson_default_options": (None, [ctypes.POINTER(DavidsonOptions)]),
    "q3x_davidson_solve": (_i, [_p, _i, ctypes.POINTER(DavidsonOptions), _p, _p, ctypes.POINTER(DavidsonReport)]),
    "q3x_davidson_solve_device": (_i, [_p, _i, ctypes.POINTER(DavidsonOptions), _p, _p,
                                       ctypes.POINTER(DavidsonReport)]),
    "q3x_server_create": (_i, [_i, _i, _d, _pp]),
    "q3x_server_destroy": (_i, [_p]),
    "q3x_server_cancel_pending": (_i, [_p, _p]),
    "q3x_server_estimate": (_i, [_p, ctypes.POINTER(BandRequest), ctypes.POINTER(_d)]),
    "q3x_memory_estimate_default_inputs": (None, [ctypes.POINTER(MemoryEstimateInputs)]),
    "q3x_memory_estimate_compute": (_i, [ctypes.POINTER(MemoryEstimateInputs), ctypes.POINTER(MemoryEstimate)]),
    "q3x_band_request_footprint": (_i, [ctypes.POINTER(BandRequest), _i, ctypes.POINTER(MemoryEstimate),
                                        ctypes.POINTER(MemoryTerm), _p, ctypes.POINTER(_d)]),
    "q3x_server_submit": (_i, [_p, ctypes.POINTER(BandRequest), ctypes.POINTER(_l)]),
    "q3x_server_wait": (_i, [_p, _l, _p, ctypes.POINTER(RequestStats)]),
    "q3x_server_in_flight_bytes": (_d, [_p]),
    "q3x_batch_create": (_i, [_i, _i, _pp]),
    "q3x_batch_destroy": (_i, [_p]),
    "q3x_batch_add": (_i, [_p, _p, _p, _p, _l, _p, _i, _i, _p, _p, ctypes.POINTER(_i)]),
    "q3x_batch_size": (_i, [_p]),
    "q3x_batch_solve": (_i, [_p, _i, ctypes.POINTER(DavidsonOptions), _pp, _p, ctypes.POINTER(DavidsonReport)]),
}

_lib = None


class Q3XError(RuntimeError):
    """A q3x call returned a non-zero status."""

    def __init__(self, status: int, message: str):
        super().__init__(f"q3x status {status}: {message}")
        self.status = status


class StateError(Q3XError):
    """Call order violated (the reference throws std::logic_error here)."""


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(f"{LIB_PATH} is not built; run `python -m q3_amd.build` "
                              "(the HIP library is required; there is no fallback)")
        L = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def check(status: int):
    if status != 0:
        msg = lib().q3x_last_error().decode(errors="replace")
        if status == 3:
            raise StateError(status, msg)
        raise Q3XError(status, msg)
