// Footprint estimate of a plane-wave run from its dimensions (PlanewaveMemoryEstimate::compute,
// iterator/kohn_sham/planewave/internal/memory_estimate.hpp:75-215).  The labels, formulas and the
// unsigned 64-bit arithmetic (including the order of the integer division in the D-block term) are
// the reference's, so an estimate printed by either agrees to the byte.  Terms of zero bytes are
// left out; the setup peak is the steady state less the solver's share plus the largest spike, and
// the peak is the taller of the two phases.
#include <cstdint>
#include <cstring>

#include "q3x_internal.h"

namespace {

using u64 = std::uint64_t;

struct Terms {
    q3x_memory_estimate* e;
    void steady(const char* label, const char* formula, u64 bytes, bool solver = false) {
        if (bytes == 0 || e->n_steady_state >= 16) return;
        e->steady_state[e->n_steady_state++] = {label, formula, static_cast<unsigned long long>(bytes)};
        e->steady_state_bytes += bytes;
        if (solver) e->solver_bytes += bytes;
    }
    void transient(const char* label, const char* formula, u64 bytes) {
        if (bytes == 0 || e->n_transient >= 4) return;
        e->transient[e->n_transient++] = {label, formula, static_cast<unsigned long long>(bytes)};
    }
};

}  // namespace

extern "C" {

void q3x_memory_estimate_default_inputs(q3x_memory_estimate_inputs* in) {
    if (!in) return;
    std::memset(in, 0, sizeof(*in));
    in->number_of_spin_components = 1;
    in->local_number_of_kpoint_states = 1;
    in->davidson_basis_expansion_factor = 2;
    in->overlap_is_identity = 0;
    in->diagonalizer_is_davidson = 1;
}

q3x_status q3x_memory_estimate_compute(const q3x_memory_estimate_inputs* in, q3x_memory_estimate* out) {
    if (!in || !out) {
        q3x::set_error("q3x_memory_estimate_compute: null argument");
        return Q3X_ERR_ARGUMENT;
    }
    std::memset(out, 0, sizeof(*out));
    const u64 c16 = 16, r8 = 8;
    const u64 nb = static_cast<u64>(in->number_of_bands);
    const u64 npw = static_cast<u64>(in->max_plane_waves_per_kpoint);
    const u64 ngm = static_cast<u64>(in->number_of_g_vectors);
    const u64 nkb = static_cast<u64>(in->number_of_projectors);
    const u64 nspin = static_cast<u64>(in->number_of_spin_components);
    const u64 nrho = static_cast<u64>(in->fft_rho_grid_points);
    const u64 nwave = static_cast<u64>(in->fft_wave_grid_points);

    // the basis: the expansion factor times the global band count, clamped by the plane-wave
    // count but never below the band count
    const int expanded = in->davidson_basis_expansion_factor * in->number_of_bands;
    const int clamped = expanded < in->max_plane_waves_per_kpoint ? expanded : in->max_plane_waves_per_kpoint;
    out->maximum_basis_size = clamped > in->number_of_bands ? clamped : in->number_of_bands;
    const u64 m = static_cast<u64>(out->maximum_basis_size);

    Terms t{out};
    if (in->diagonalizer_is_davidson) {
        t.steady("Davidson basis V, HV", "2 x M x npw x 16", 2 * m * npw * c16, true);
        if (!in->overlap_is_identity) t.steady("Davidson basis SV", "M x npw x 16", m * npw * c16, true);
        t.steady("rotation scratch", "nb x npw x 16", nb * npw * c16, true);
        if (!in->overlap_is_identity) t.steady("Ritz overlap vectors", "nb x npw x 16", nb * npw * c16, true);
        t.steady("subspace matrices", "4 x M^2 x 16", 4 * m * m * c16, true);
    }
    t.steady("wavefunctions", "nkstates x nb x npw x 16",
             static_cast<u64>(in->local_number_of_kpoint_states) * nb * npw * c16);
    t.steady("nonlocal projectors", "nkb x npw x 16", nkb * npw * c16);
    if (in->number_of_atoms > 0)
        t.steady("block-diagonal D matrix", "nkb^2 / nat x 16",
                 nkb * nkb / static_cast<u64>(in->number_of_atoms) * c16);
    t.steady("Broyden history", "3 x ndim x nspin x ngm x 16",
             3 * static_cast<u64>(in->mixing_history_depth) * nspin * ngm * c16);
    t.steady("G-vector table", "ngm x 80", ngm * 80);
    t.steady("FFT buffers", "2 x (Nrho + 2 x Nwave) x 16", 2 * (nrho + 2 * nwave) * c16);
    t.steady("density and potential", "(4 + 3 x nspin) x Nrho x 8", (4 + 3 * nspin) * nrho * r8);

    const u64 cols = static_cast<u64>(in->number_of_atomic_orbitals > in->number_of_bands ? in->number_of_atomic_orbitals
                                                                                          : in->number_of_bands);
    t.transient("wavefunction initialization", "5 x max(natwfc, nb) x npw x 16", 5 * cols * npw * c16);

    u64 largest = 0;
    for (int i = 0; i < out->n_transient; ++i)
        if (out->transient[i].bytes > largest) largest = out->transient[i].bytes;
    out->setup_peak_bytes = out->steady_state_bytes - out->solver_bytes + largest;
    out->peak_bytes = out->steady_state_bytes > out->setup_peak_bytes ? out->steady_state_bytes : out->setup_peak_bytes;
    return Q3X_OK;
}

}  // extern "C"
