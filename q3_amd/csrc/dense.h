// FP64 dense Hermitian linear algebra for the Davidson subspace (host, OpenMP).
#pragma once
#include <complex>

namespace q3x {

using zc = std::complex<double>;

// A = U^H U (upper).  Returns 0, or j+1 if the leading minor j+1 is not positive definite.
int chol_upper(int n, zc* A, int lda);
// U <- U^{-1} for upper-triangular U (lower part ignored and zeroed).
void tri_inv_upper(int n, zc* U, int ldu);
// Generalized Hermitian eigenproblem H x = w S x (upper triangles of H and S are read, as
// LAPACK zhegv itype 1 'V' 'U').  All n eigenvalues ascending into w; eigenvectors of the
// lowest nvec into X (n x nvec, ld ldx), normalised X^H S X = I.  Returns 0 or an info code
// (n + j for an S that is not positive definite, like zhegv).
int hegv(int n, const zc* H, int ldh, const zc* S, int lds, double* w, zc* X, int ldx, int nvec);

}  // namespace q3x
