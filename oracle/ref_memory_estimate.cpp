// Test infrastructure: a C entry point around the reference's own PlanewaveMemoryEstimate::compute
// (src/layer_4_computation/iterator/kohn_sham/planewave/internal/memory_estimate.hpp, a header with
// no dependencies beyond the standard library), compiled from the reference's source tree into
// oracle/_ref/ by `make -C oracle ref`.  Only the parity tests load it, as the checker of the
// library's q3x_memory_estimate_compute.  Nothing of the reference is copied here: the header is
// included from where it lies.
#include <cstdint>
#include <cstring>

#include "layer_4_computation/iterator/kohn_sham/planewave/internal/memory_estimate.hpp"

namespace {
void put(char* dst, const std::string& s) {
    std::strncpy(dst, s.c_str(), 95);
    dst[95] = '\0';
}
}  // namespace

// in: the 14 input fields in declaration order as int64 (the two booleans as 0/1).
// Outputs: totals[4] = steady, solver, setup peak, peak; per term 96-byte label and formula
// slots; returns the maximum basis size.
extern "C" int ref_memory_estimate(const std::int64_t* in, unsigned long long* totals, int* n_steady,
                                   char* steady_labels, char* steady_formulas, unsigned long long* steady_bytes,
                                   int* n_transient, char* transient_labels, char* transient_formulas,
                                   unsigned long long* transient_bytes) {
    q3::PlanewaveMemoryEstimateInputs x;
    x.number_of_bands = static_cast<int>(in[0]);
    x.max_plane_waves_per_kpoint = static_cast<int>(in[1]);
    x.number_of_g_vectors = static_cast<int>(in[2]);
    x.number_of_projectors = static_cast<int>(in[3]);
    x.number_of_atomic_orbitals = static_cast<int>(in[4]);
    x.number_of_atoms = static_cast<int>(in[5]);
    x.number_of_spin_components = static_cast<int>(in[6]);
    x.local_number_of_kpoint_states = static_cast<int>(in[7]);
    x.mixing_history_depth = static_cast<int>(in[8]);
    x.fft_rho_grid_points = in[9];
    x.fft_wave_grid_points = in[10];
    x.davidson_basis_expansion_factor = static_cast<int>(in[11]);
    x.overlap_is_identity = in[12] != 0;
    x.diagonalizer_is_davidson = in[13] != 0;
    const q3::PlanewaveMemoryEstimate e = q3::PlanewaveMemoryEstimate::compute(x);
    totals[0] = e.steady_state_bytes;
    totals[1] = e.solver_bytes;
    totals[2] = e.setup_peak_bytes;
    totals[3] = e.peak_bytes;
    *n_steady = static_cast<int>(e.steady_state.size());
    for (size_t i = 0; i < e.steady_state.size() && i < 16; ++i) {
        put(steady_labels + 96 * i, e.steady_state[i].label);
        put(steady_formulas + 96 * i, e.steady_state[i].formula);
        steady_bytes[i] = e.steady_state[i].bytes;
    }
    *n_transient = static_cast<int>(e.transient.size());
    for (size_t i = 0; i < e.transient.size() && i < 4; ++i) {
        put(transient_labels + 96 * i, e.transient[i].label);
        put(transient_formulas + 96 * i, e.transient[i].formula);
        transient_bytes[i] = e.transient[i].bytes;
    }
    return e.maximum_basis_size;
}

// the reference's default inputs (PlanewaveMemoryEstimateInputs{}), same order as above
extern "C" void ref_memory_estimate_defaults(std::int64_t* out) {
    const q3::PlanewaveMemoryEstimateInputs x{};
    const std::int64_t v[14] = {x.number_of_bands, x.max_plane_waves_per_kpoint, x.number_of_g_vectors,
                                x.number_of_projectors, x.number_of_atomic_orbitals, x.number_of_atoms,
                                x.number_of_spin_components, x.local_number_of_kpoint_states,
                                x.mixing_history_depth, x.fft_rho_grid_points, x.fft_wave_grid_points,
                                x.davidson_basis_expansion_factor, x.overlap_is_identity ? 1 : 0,
                                x.diagonalizer_is_davidson ? 1 : 0};
    std::memcpy(out, v, sizeof(v));
}
