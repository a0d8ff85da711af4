// Gamma-trick kernels (gamma_kernels.hip).  Leading dimensions are in complex elements.
#pragma once
#include <hip/hip_runtime.h>

namespace q3x {

// phi (extended list: npw half-sphere entries, then the npw - 1 mirrors; (nvec + 1) / 2 columns)
// from the half-sphere bands psi (npw x nvec)
hipError_t launch_gamma_pack(int npw, int nvec, const double* psi, long ldpsi, double* phi, long ldphi,
                             hipStream_t st);
// H psi (npw x nvec) from the transformed pairs f (extended list, (nvec + 1) / 2 columns)
hipError_t launch_gamma_unpack(int npw, int nvec, const double* f, long ldf, double* hpsi, long ldh, hipStream_t st);
// C (m x n) = 2 Re(C) - Re(conj(A[0, i]) B[0, j]) + 0i: the full-sphere real product of half-sphere
// columns from their half-sphere product C = A^H B
hipError_t launch_gamma_gram_fix(int m, int n, const double* A, long lda, const double* B, long ldb, double* C,
                                 long ldc, hipStream_t st);
// the same for projector overlaps P = beta^H psi (P and beta complex64 when fp32)
hipError_t launch_gamma_overlap_fix(bool fp32, int np, int nvec, const void* beta, long ldb, const double* psi,
                                    long ldpsi, void* P, long ldp, hipStream_t st);
// X[0, c] <- Re X[0, c] (enforce_gamma_g0_real)
hipError_t launch_g0_real(int ncol, double* X, long ldx, hipStream_t st);
// n2[c] <- 2 n2[c] - |X[0, c]|^2 (half-sphere squared norms to full-sphere ones)
hipError_t launch_gamma_norm_fix(int ncol, const double* X, long ldx, double* n2, hipStream_t st);
// imaginary parts of X (rows x ncol) to zero
hipError_t launch_zero_imag(int rows, int ncol, double* X, long ldx, hipStream_t st);

}  // namespace q3x
