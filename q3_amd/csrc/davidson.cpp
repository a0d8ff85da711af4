// Block Davidson driver (placeholder until the device implementation lands).
#include "runtime.h"
extern "C" void q3x_davidson_default_options(q3x_davidson_options* o) {
    o->tolerance_ry = 1e-2; o->max_iterations = 20; o->basis_factor = 2; o->n_occupied = 0;
    o->precision = Q3X_FP64; o->seed_from_psi = 0; o->final_fp64_sweep = 1;
}
extern "C" q3x_status q3x_davidson_solve(q3x_hamiltonian, int, const q3x_davidson_options*, double*, double*,
                                         q3x_davidson_report*) {
    q3x::set_error("q3x_davidson_solve: not built yet");
    return Q3X_ERR_INTERNAL;
}
