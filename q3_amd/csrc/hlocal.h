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
    const void* twp;              // exp(-2 pi i j/n1), j < n1, for the two-level plane kernel
    const void* twp3;             // exp(-2 pi i j/n3), j < n3, for the two-level z-stick kernels
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

// Two-level plane kernel (plane_tl.hip) for square planes n1 = n2 of the sizes it is built for.
struct PlaneTlConfig {
    bool ok = false;
    int lc = 0, lr = 0;  // LDS strides (complex elements) of a column and of a row in flight
    int rb = 0;          // rows in flight per wave
    int row_waves = 0;   // waves with a row buffer (the first row_waves of threads / 64)
    int threads = 0;
    int workgroups = 0;  // persistent workgroups of a launch (each loops over planes)
    size_t lds = 0;
    bool plane_global = false;  // plane too large for LDS: kept in a global slot per workgroup
    size_t plane_bytes = 0;     // bytes of one plane slot
};
bool plane_tl_supported(int n1, int n2);
PlaneTlConfig plane_tl_config(int n, int ncol, bool fp32, size_t max_lds, int num_cus);
// Two-level z-stick kernels (plane_tl.hip) for the n3 the plane sizes list holds.
struct StickTlConfig {
    bool ok = false;
    int spb = 0;    // sticks per workgroup of the inverse pass
    int spb_f = 0;  // and of the forward pass (the stick buffer's layout does not depend on it)
    int lz = 0;     // LDS stride of a stick (complex elements)
    int threads = 0;
    size_t lds = 0, lds_f = 0;
};
bool stick_tl_supported(int n3);
StickTlConfig stick_tl_config(int n3, bool fp32);
// dir > 0: zsticks inverse (psi -> Z); dir < 0: forward + gather + kinetic (Z -> hpsi)
template <typename T>
hipError_t launch_sticks_tl(const HLocalArgs& a, int dir, const double* psi, long ldpsi, cx<T>* zbuf, double* hpsi,
                            long ldh, const StickTlConfig& c, int nvec, hipStream_t st);
// pglob: nvec * n3 plane slots of c.plane_bytes when c.plane_global, else unused
template <typename T>
hipError_t launch_plane_tl(const HLocalArgs& a, const T* vplane, cx<T>* zbuf, const PlaneTlConfig& c, int nvec,
                           void* pglob, hipStream_t st);
// Charge density of nvec columns of the stick buffer (after the inverse z-stick pass):
// rho(i1, i2, i3) += wts[2c] Re(f_c(r))^2 + wts[2c+1] Im(f_c(r))^2, rho C-order (i1, i2, i3), FP64.
// Needs the two-level plane configuration with the plane in LDS.
hipError_t launch_plane_density(const HLocalArgs& a, const cx<double>* zbuf, const PlaneTlConfig& c, int nvec,
                                const double* wts, double* rho, hipStream_t st);
// psi / hpsi: complex128, column per band with leading dimension ld (in complex elements).
// marks (optional): 4 events recorded before/between/after the three kernels.
// vplane: V(r)/N in plane-major layout [i3][i2][i1], double or float to match fp32.
// zbuf: nvec * n3 * nst complex elements of the compute type.
// tl: configuration of the two-level plane kernel, or null (or !ok) for the generic one.
hipError_t launch_hlocal(const HLocalArgs& a, bool fp32, const void* vplane, void* zbuf, const double* psi,
                         long ldpsi, double* hpsi, long ldh, int nvec, hipStream_t st,
                         hipEvent_t* marks = nullptr, const PlaneTlConfig* tl = nullptr, void* pglob = nullptr,
                         const StickTlConfig* tz = nullptr);

}  // namespace q3x
