// Complex GEMM on MFMA (zgemm.hip).
#pragma once
#include <hip/hip_runtime.h>

#include <cstddef>

namespace q3x {

// Element types of A, B, the MFMA/accumulation type, and C.  D = complex128, F = complex64.
enum class GemmTypes { D_D_D_D, F_D_F_F, F_F_F_D, D_D_F_D, F_F_F_F };

// C (m x n) = alpha * op(A) * B + beta * C ; opa = 0: A is m x k ; opa = 1: A is k x m, op = A^H.
// Leading dimensions in complex elements.  work (may be null) enables split-K for skinny shapes.
hipError_t zgemm(GemmTypes types, int opa, int m, int n, int k, double alpha, const void* A, long lda,
                 const void* B, long ldb, double beta, void* C, long ldc, void* work, size_t work_bytes,
                 hipStream_t st);

}  // namespace q3x
