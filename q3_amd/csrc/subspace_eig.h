// Generalized Hermitian eigensolver for the Davidson subspace problem, on the device
// (subspace_eig.hip).  Stands in for GPUEigensolver::generalized_hermitian_eigenvalue_solve
// (math/linear_algebra/gpu/lapack/init.hpp:77-82) with zhegv itype-1 semantics.
#pragma once
#include <hip/hip_runtime.h>

#include <cstddef>

#include "runtime.h"

namespace q3x {

// Device workspace for subspace_eig(n, nvec, ...), in bytes.
size_t subspace_eig_workspace(int n, int nvec);

struct SubspaceEigStats {
    int newton_schulz_iterations = 0;  // 0: first-order Loewdin factor was accurate enough
    double s_deviation = 0.0;          // max |S - I| over the compacted matrix
    // solves redone on one workgroup because a grid barrier of the multi-workgroup
    // tridiagonalisation gave up (added to, never reset, by subspace_eig)
    int barrier_redos = 0;
};

// Test hook: the next `count` multi-workgroup tridiagonalisations start with their give-up flag
// set, so each is redone on one workgroup.
void force_barrier_giveups(int count);
// Test hook: bit i of mask makes the i-th subspace_eig call from now return n + 1 without solving.
void fail_eigensolves(unsigned long long mask);
// Redos on one workgroup since the process started.
long barrier_redos_total();

// Solves H x = w S x for the nvec lowest pairs.  H and S (device, complex128, column-major,
// leading dimensions ldh / lds) are read from their upper triangles, as zhegv 'U' does.
// On success w_host holds the eigenvalues (ascending) and X (device, ld ldx) the
// S-orthonormal eigenvectors; returns 0.  Returns n + 1 when S is not numerically positive
// definite (the zhegv info > n case) and -1 on a device error (*err set).
// work: subspace_eig_workspace(n, nvec) bytes; gemm_work: split-K workspace of zgemm.
// num_cus: compute units of the device (bounds the workgroups of the multi-workgroup reduction).
// pin (optional): pinned staging for the copies of eigenvalues and status words to the host.
// lowdin_bound: the first-order factor S^(-1/2) ~ 1.5 I - 0.5 S is used when n max|S - I| is at
// most this (its error is about (3/8) (n max|S - I|)^2), else Newton-Schulz; 1e-7 keeps the
// factor to rounding, callers whose S carries complex64 rounding pass a looser bound.
int subspace_eig(int n, const double* H, long ldh, const double* S, long lds, int nvec, double* w_host, double* X,
                 long ldx, void* work, size_t work_bytes, void* gemm_work, size_t gemm_bytes, size_t max_lds,
                 int num_cus, hipStream_t st, hipError_t* err, SubspaceEigStats* stats, Profiler* prof = nullptr,
                 HostStage* pin = nullptr, double lowdin_bound = 1e-7);

}  // namespace q3x
