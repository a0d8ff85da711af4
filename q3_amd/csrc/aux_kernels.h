// Small kernels around the H-apply and Davidson iteration (aux_kernels.hip).
#pragma once
#include <hip/hip_runtime.h>

namespace q3x {

hipError_t launch_dblock_apply(bool fp32, int nproj, int nvec, const int* pstart, const int* psize, const long* poff,
                               const double* D, const void* P, void* W, hipStream_t st);
hipError_t launch_teter(int npw, int ncol, const double* ekin, const double* eps, double* R, long ldr,
                        hipStream_t st);
hipError_t launch_column_norm2(int npw, int ncol, const double* X, long ldx, double* out, hipStream_t st);
hipError_t launch_scale_copy(int npw, int ncol, const double* X, long ldx, const double* s, double* Y, long ldy,
                             hipStream_t st);
// Y = a_c X + beta Y per column (a null -> 1)
hipError_t launch_axpby(int npw, int ncol, const double* a, const double* X, long ldx, double beta, double* Y,
                        long ldy, hipStream_t st);
hipError_t launch_zero(double* p, long ndouble, hipStream_t st);
// Y[:, c] = s[c] X[:, src[c]] (s null -> 1), complex columns of length rows
// R[:, c] = Teter(HV[:, src[c]] - eps[c] V[:, src[c]]) (complex columns, leading dims in complex)
hipError_t launch_residual_columns_teter(int npw, int ncol, const int* src, const double* eps, const double* ekin,
                                         const double* V, long ldv, const double* HV, long ldh, double* R, long ldr,
                                         hipStream_t st);
hipError_t launch_gather_columns(int rows, int ncol, const int* src, const double* s, const double* X, long ldx,
                                 double* Y, long ldy, hipStream_t st);
// plain copy of n16 16-byte elements (HBM byte-counter reference)
hipError_t launch_copy16(const void* src, void* dst, long n16, hipStream_t st);
hipError_t launch_unit_columns(const int* idx, int ncol, double* V, long ldv, hipStream_t st);
// X (n x n, leading dimension ldx, complex128) = identity
hipError_t launch_identity(int n, double* X, long ldx, hipStream_t st);

}  // namespace q3x
