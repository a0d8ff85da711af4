// Launch interface of the fused local-Hamiltonian kernels (hlocal.hip).
#pragma once
#include <hip/hip_runtime.h>

#include "fft_lds.h"

namespace q3x {

struct HLocalArgs {
    FftPlan p1, p2, p3;           // FFT lengths along i1, i2, i3
    const void* tw1;              // twiddle tables exp(-2 pi i k/n), element type = compute type
    const void* tw2;
    const void* tw3;
    int nst;                      // occupied z-sticks
    int ncol;                     // occupied i1 columns
    int sticks_per_block;
    int rows_per_pass;
    const int* stick_off;         // nst + 1
    const int* stick_i2;          // nst
    const int* stick_col;         // nst
    const int* col_i1;            // ncol
    const int* ent_pw;            // entries (stick order) -> plane-wave index
    const int* ent_pos;           // entries -> stick * n3 + i3
    const double* ekin;           // npw, Ry
    const int* outside;           // plane waves outside the box (kinetic only)
    int n_outside;
};

size_t hlocal_lds_bytes(const HLocalArgs& a, bool fp32, int which);
// psi / hpsi: complex128, column per band with leading dimension ld (in complex elements).
// marks (optional): 4 events recorded before/between/after the three kernels.
// vplane: V(r)/N in plane-major layout [i3][i2][i1], double or float to match fp32.
// zbuf: nvec * n3 * nst complex elements of the compute type.
hipError_t launch_hlocal(const HLocalArgs& a, bool fp32, const void* vplane, void* zbuf, const double* psi,
                         long ldpsi, double* hpsi, long ldh, int nvec, hipStream_t st,
                         hipEvent_t* marks = nullptr);

}  // namespace q3x
