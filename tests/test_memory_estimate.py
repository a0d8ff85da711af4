"""The footprint estimate (q3x_memory_estimate_compute, PlanewaveMemoryEstimate::compute,
iterator/kohn_sham/planewave/internal/memory_estimate.hpp:75-215) and the server's admission
footprint built on it.

* The reference's own unit tests, restated on the library
  (tests/unit/.../planewave/internal/test_memory_estimate.cpp): the 432-atom gamma-only silicon
  deck's basis sizes, the per-band-and-plane-wave ratios 6 and 10, clamping, the 35-50 GiB
  window, the phases, the direct solver, the identity overlap.
* The reference's header itself, compiled from the reference tree into oracle/_ref/ (skipped where
  that tree is absent): every term, label, formula and total equal, on the deck above, its
  variants and 400 random inputs, and the defaults equal.
* A request's admission footprint is the reference estimate of its dimensions -- the
  wavefunctions once, the solver, projector and D lines per running k-point -- plus the itemised
  device terms; a k-path request counts one running k-point.
"""
import ctypes
import os

import numpy as np
import pytest

from q3_amd import BandRequestSpec, PlanewaveBasis
from q3_amd import workloads as W
from q3_amd.planewave import band_request_footprint, memory_estimate

GIB = 1073741824.0
REF = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "oracle", "_ref",
                   "libref_memory_estimate.so")
FIELDS = ("number_of_bands", "max_plane_waves_per_kpoint", "number_of_g_vectors", "number_of_projectors",
          "number_of_atomic_orbitals", "number_of_atoms", "number_of_spin_components",
          "local_number_of_kpoint_states", "mixing_history_depth", "fft_rho_grid_points", "fft_wave_grid_points",
          "davidson_basis_expansion_factor", "overlap_is_identity", "diagonalizer_is_davidson")


def large_silicon(**over):
    """The reference test's 432-atom gamma-only silicon deck (test_memory_estimate.cpp:13-30)."""
    x = dict(davidson_basis_expansion_factor=4, number_of_bands=1040, max_plane_waves_per_kpoint=167155,
             number_of_g_vectors=1338865, number_of_projectors=7776, number_of_atomic_orbitals=1728,
             number_of_atoms=432, number_of_spin_components=1, local_number_of_kpoint_states=1,
             mixing_history_depth=8, fft_rho_grid_points=2560000, fft_wave_grid_points=2560000,
             overlap_is_identity=1)
    x.update(over)
    return x


def per_plane_wave(e):
    return sum(b for _l, f, b in e["steady_state"] if "npw" in f and "nkb" not in f)


def test_reference_known_answers():
    e = memory_estimate(**large_silicon())
    assert e["maximum_basis_size"] == 4160
    unit = 1040 * 167155 * 16
    assert abs(per_plane_wave(e) / unit - 10.0) < 0.05
    assert 35.0 < e["steady_state_bytes"] / GIB < 50.0
    assert all("collapse" not in label for label, _f, _b in e["steady_state"])
    assert e["solver_bytes"] * 2 > e["steady_state_bytes"]
    assert e["peak_bytes"] == max(e["steady_state_bytes"], e["setup_peak_bytes"])
    assert e["peak_bytes"] < e["steady_state_bytes"] + e["setup_peak_bytes"]
    d = memory_estimate(**large_silicon(davidson_basis_expansion_factor=2))
    assert d["maximum_basis_size"] == 2080 and abs(per_plane_wave(d) / unit - 6.0) < 0.05
    assert memory_estimate(**large_silicon(number_of_bands=100, max_plane_waves_per_kpoint=250))[
        "maximum_basis_size"] == 250
    assert memory_estimate(**large_silicon(number_of_bands=100, max_plane_waves_per_kpoint=50))[
        "maximum_basis_size"] == 100
    direct = memory_estimate(**large_silicon(diagonalizer_is_davidson=0))
    assert direct["solver_bytes"] == 0 and direct["peak_bytes"] > direct["steady_state_bytes"]
    us = memory_estimate(**large_silicon(overlap_is_identity=0))
    assert all("SV" not in l and "Ritz overlap" not in l for l, _f, _b in e["steady_state"])
    assert us["steady_state_bytes"] - e["steady_state_bytes"] == (4160 + 1040) * 167155 * 16


def _ref_lib():
    if not os.path.exists(REF):
        pytest.skip("oracle/_ref not built (the reference tree is absent)")
    L = ctypes.CDLL(REF)
    L.ref_memory_estimate.restype = ctypes.c_int
    return L


def _ref(L, x):
    d = np.zeros(14, np.int64)
    L.ref_memory_estimate_defaults(d.ctypes.data_as(ctypes.c_void_p))
    vals = dict(zip(FIELDS, (int(v) for v in d)))
    vals.update(x)
    inp = np.array([vals[f] for f in FIELDS], np.int64)
    tot = np.zeros(4, np.uint64)
    ns, nt = ctypes.c_int(), ctypes.c_int()
    sl, sf, tl, tf = (ctypes.create_string_buffer(96 * 16) for _ in range(4))
    sb, tb = np.zeros(16, np.uint64), np.zeros(4, np.uint64)
    m = L.ref_memory_estimate(inp.ctypes.data_as(ctypes.c_void_p), tot.ctypes.data_as(ctypes.c_void_p),
                              ctypes.byref(ns), sl, sf, sb.ctypes.data_as(ctypes.c_void_p), ctypes.byref(nt), tl, tf,
                              tb.ctypes.data_as(ctypes.c_void_p))

    def strs(buf, n):
        return [buf.raw[96 * i:96 * (i + 1)].split(b"\0")[0].decode() for i in range(n)]
    steady = list(zip(strs(sl, ns.value), strs(sf, ns.value), (int(b) for b in sb[:ns.value])))
    trans = list(zip(strs(tl, nt.value), strs(tf, nt.value), (int(b) for b in tb[:nt.value])))
    return {"steady_state": steady, "transient": trans, "steady_state_bytes": int(tot[0]),
            "solver_bytes": int(tot[1]), "setup_peak_bytes": int(tot[2]), "peak_bytes": int(tot[3]),
            "maximum_basis_size": m}, vals


def test_defaults_equal_the_reference():
    L = _ref_lib()
    d = np.zeros(14, np.int64)
    L.ref_memory_estimate_defaults(d.ctypes.data_as(ctypes.c_void_p))
    from q3_amd import _lib
    from q3_amd._lib import lib
    x = _lib.MemoryEstimateInputs()
    lib().q3x_memory_estimate_default_inputs(ctypes.byref(x))
    assert [int(getattr(x, f)) for f in FIELDS] == [int(v) for v in d]


def test_equals_the_reference_header_on_the_deck_and_random_inputs():
    L = _ref_lib()
    cases = [large_silicon(), large_silicon(davidson_basis_expansion_factor=2),
             large_silicon(number_of_bands=100, max_plane_waves_per_kpoint=250),
             large_silicon(overlap_is_identity=0), large_silicon(diagonalizer_is_davidson=0),
             large_silicon(number_of_atoms=0), {}]
    rng = np.random.default_rng(11)
    for _ in range(400):
        cases.append({
            "number_of_bands": int(rng.integers(1, 5000)), "max_plane_waves_per_kpoint": int(rng.integers(1, 400000)),
            "number_of_g_vectors": int(rng.integers(0, 3_000_000)), "number_of_projectors": int(rng.integers(0, 20000)),
            "number_of_atomic_orbitals": int(rng.integers(0, 6000)), "number_of_atoms": int(rng.integers(0, 1000)),
            "number_of_spin_components": int(rng.integers(1, 3)),
            "local_number_of_kpoint_states": int(rng.integers(0, 64)), "mixing_history_depth": int(rng.integers(0, 12)),
            "fft_rho_grid_points": int(rng.integers(0, 10_000_000)), "fft_wave_grid_points": int(rng.integers(0, 5_000_000)),
            "davidson_basis_expansion_factor": int(rng.integers(1, 6)), "overlap_is_identity": int(rng.integers(0, 2)),
            "diagonalizer_is_davidson": int(rng.integers(0, 2))})
    for x in cases:
        ref, vals = _ref(L, x)
        assert memory_estimate(**vals) == ref, x


def _serving_spec(nat, nk, kpath):
    r = next(q for q in W.cfg5_requests(200, seed=4) if len(q.crystal.species) == nat)
    sp = W.serving_species()
    basis = PlanewaveBasis(r.crystal, r.ecutwfc_ry, 0.0, True, (1, 1, 1))
    bs, db = W.d_blocks_for(r.crystal, sp)
    V = np.zeros(int(np.prod(basis.wave_dims)))
    return BandRequestSpec(basis, V, r.kpoints[:nk], sp, bs, db, r.nbands, r.noccupied, kpath_propagation=kpath,
                           band_path_segments=1 if kpath else 0, band_path_points_per_segment=nk - 1 if kpath else 0)


@pytest.mark.parametrize("nat,nk,kpath,concurrent", [(2, 12, 0, 8), (8, 3, 0, 8), (16, 20, 1, 8), (2, 5, 0, 2)])
def test_request_footprint_is_the_reference_estimate_plus_device_terms(nat, nk, kpath, concurrent):
    spec = _serving_spec(nat, nk, kpath)
    f = band_request_footprint(spec, 1 if kpath else concurrent)
    ref = f["reference"]
    terms = {label: b for label, _f, b in ref["steady_state"]}
    held = ("Davidson basis V, HV", "rotation scratch", "subspace matrices", "nonlocal projectors",
            "block-diagonal D matrix")
    running = 1 if kpath else min(nk, concurrent)
    per_k = sum(terms.get(h, 0) for h in held) + sum(b for _l, _f, b in f["device_terms"])
    assert f["bytes"] == pytest.approx(terms["wavefunctions"] + per_k * running, rel=1e-12)
    # the reference part is the estimate of the request's own dimensions
    npw = spec.basis.compute_gk(spec.kpoints[0]).npw
    nproj = int(np.sum(spec.block_sizes))
    again = memory_estimate(number_of_bands=spec.nbands, max_plane_waves_per_kpoint=int(npw * 1.05) + 16,
                            number_of_g_vectors=spec.basis.number_of_g_vectors, number_of_projectors=nproj,
                            number_of_atoms=len(spec.basis.crystal.species), local_number_of_kpoint_states=nk,
                            fft_rho_grid_points=int(np.prod(spec.basis.rho_dims)),
                            fft_wave_grid_points=int(np.prod(spec.basis.wave_dims)), overlap_is_identity=1)
    assert again == ref
    assert terms["wavefunctions"] == nk * spec.nbands * (int(npw * 1.05) + 16) * 16
