// Gamma-trick kernels (hamiltonian_applicator/serial.hpp:283-418, 519-566 and the gamma paths
// of subspace_manager/serial.hpp).  A Gamma-trick band holds the half-sphere coefficients of a
// real function (psi(-G) = conj(psi(G)), G = 0 is plane wave 0 and real).  Two bands share one
// complex transform: the extended list of the applicator holds the half sphere followed by the
// mirrors -G_j of j = 1..npw-1, so mirror j sits at npw + j - 1.
#include <hip/hip_runtime.h>

#include "gamma_kernels.h"

namespace q3x {
namespace {

// phi[:, p] (extended list) from bands 2p, 2p+1 of psi: psi_1 + i psi_2 at +G, conj(psi_1) +
// i conj(psi_2) at -G, (Re psi_1(0), Re psi_2(0)) at G = 0; an odd last band alone: psi_1 at +G
// (G = 0 as it is), conj(psi_1) at -G.
__global__ void gamma_pack(int npw, int nvec, const double* __restrict__ psi, long ldpsi, double* __restrict__ phi,
                           long ldphi) {
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    const int p = blockIdx.y;
    if (j >= npw) return;
    const int b1 = 2 * p, b2 = 2 * p + 1;
    const double* p1 = psi + 2L * b1 * ldpsi;
    double* out = phi + 2L * p * ldphi;
    const double r1 = p1[2 * j], i1 = p1[2 * j + 1];
    if (b2 < nvec) {
        const double* p2 = psi + 2L * b2 * ldpsi;
        const double r2 = p2[2 * j], i2 = p2[2 * j + 1];
        if (j == 0) {
            out[0] = r1;
            out[1] = r2;
            return;
        }
        out[2 * j] = r1 - i2;
        out[2 * j + 1] = i1 + r2;
        const long m = npw + j - 1;
        out[2 * m] = r1 + i2;
        out[2 * m + 1] = -i1 + r2;
    } else {
        out[2 * j] = r1;
        out[2 * j + 1] = i1;
        if (j == 0) return;
        const long m = npw + j - 1;
        out[2 * m] = r1;
        out[2 * m + 1] = -i1;
    }
}

// Separation after the forward transform (the kinetic term rides along: T(-G) = T(G)):
// h_1 = (F(G) + conj F(-G)) / 2, h_2 = -i (F(G) - conj F(-G)) / 2, G = 0: Re / Im of F(0);
// an odd last band: h_1 = F(G).
__global__ void gamma_unpack(int npw, int nvec, const double* __restrict__ f, long ldf, double* __restrict__ hpsi,
                             long ldh) {
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    const int p = blockIdx.y;
    if (j >= npw) return;
    const int b1 = 2 * p, b2 = 2 * p + 1;
    const double* in = f + 2L * p * ldf;
    double* h1 = hpsi + 2L * b1 * ldh;
    const double fr = in[2 * j], fi = in[2 * j + 1];
    if (b2 < nvec) {
        double* h2 = hpsi + 2L * b2 * ldh;
        if (j == 0) {
            h1[0] = fr;
            h1[1] = 0.0;
            h2[0] = fi;
            h2[1] = 0.0;
            return;
        }
        const long m = npw + j - 1;
        const double mr = in[2 * m], mi = -in[2 * m + 1];  // conj F(-G)
        h1[2 * j] = (fr + mr) * 0.5;
        h1[2 * j + 1] = (fi + mi) * 0.5;
        const double dr = (fr - mr) * 0.5, di = (fi - mi) * 0.5;
        h2[2 * j] = di;
        h2[2 * j + 1] = -dr;
    } else {
        h1[2 * j] = fr;
        h1[2 * j + 1] = fi;
    }
}

// C = 2 Re(C) - Re(conj(A[0, i]) B[0, j]), imaginary part 0 (complex128, leading dims complex).
__global__ void gamma_gram_fix(int m, int n, const double* __restrict__ A, long lda, const double* __restrict__ B,
                               long ldb, double* __restrict__ C, long ldc) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const int j = blockIdx.y;
    if (i >= m || j >= n) return;
    const double* a0 = A + 2L * i * lda;
    const double* b0 = B + 2L * j * ldb;
    const double g0 = a0[0] * b0[0] + a0[1] * b0[1];
    double* c = C + 2L * (static_cast<long>(j) * ldc + i);
    c[0] = 2.0 * c[0] - g0;
    c[1] = 0.0;
}

// P[p, v] = 2 Re(P) - Re(conj(beta[0, p]) psi[0, v]), imaginary part 0.
template <typename TB, typename TP>
__global__ void gamma_overlap_fix(int np, int nvec, const TB* __restrict__ beta, long ldb,
                                  const double* __restrict__ psi, long ldpsi, TP* __restrict__ P, long ldp) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    const int v = blockIdx.y;
    if (p >= np || v >= nvec) return;
    const TB* b0 = beta + 2L * p * ldb;
    const double* s0 = psi + 2L * v * ldpsi;
    const double g0 = static_cast<double>(b0[0]) * s0[0] + static_cast<double>(b0[1]) * s0[1];
    TP* c = P + 2L * (static_cast<long>(v) * ldp + p);
    c[0] = static_cast<TP>(2.0 * static_cast<double>(c[0]) - g0);
    c[1] = static_cast<TP>(0.0);
}

__global__ void g0_real(int ncol, double* __restrict__ X, long ldx) {
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c < ncol) X[2L * c * ldx + 1] = 0.0;
}

__global__ void gamma_norm_fix(int ncol, const double* __restrict__ X, long ldx, double* __restrict__ n2) {
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= ncol) return;
    const double* x0 = X + 2L * c * ldx;
    n2[c] = 2.0 * n2[c] - (x0[0] * x0[0] + x0[1] * x0[1]);
}

__global__ void zero_imag(int rows, int ncol, double* __restrict__ X, long ldx) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const int c = blockIdx.y;
    if (i < rows && c < ncol) X[2L * (static_cast<long>(c) * ldx + i) + 1] = 0.0;
}

inline unsigned blocks(long n, int t) { return static_cast<unsigned>((n + t - 1) / t); }

}  // namespace

hipError_t launch_gamma_pack(int npw, int nvec, const double* psi, long ldpsi, double* phi, long ldphi,
                             hipStream_t st) {
    if (npw <= 0 || nvec <= 0) return hipSuccess;
    hipLaunchKernelGGL(gamma_pack, dim3(blocks(npw, 256), (nvec + 1) / 2), dim3(256), 0, st, npw, nvec, psi, ldpsi,
                       phi, ldphi);
    return hipGetLastError();
}

hipError_t launch_gamma_unpack(int npw, int nvec, const double* f, long ldf, double* hpsi, long ldh, hipStream_t st) {
    if (npw <= 0 || nvec <= 0) return hipSuccess;
    hipLaunchKernelGGL(gamma_unpack, dim3(blocks(npw, 256), (nvec + 1) / 2), dim3(256), 0, st, npw, nvec, f, ldf,
                       hpsi, ldh);
    return hipGetLastError();
}

hipError_t launch_gamma_gram_fix(int m, int n, const double* A, long lda, const double* B, long ldb, double* C,
                                 long ldc, hipStream_t st) {
    if (m <= 0 || n <= 0) return hipSuccess;
    hipLaunchKernelGGL(gamma_gram_fix, dim3(blocks(m, 128), n), dim3(128), 0, st, m, n, A, lda, B, ldb, C, ldc);
    return hipGetLastError();
}

hipError_t launch_gamma_overlap_fix(bool fp32, int np, int nvec, const void* beta, long ldb, const double* psi,
                                    long ldpsi, void* P, long ldp, hipStream_t st) {
    if (np <= 0 || nvec <= 0) return hipSuccess;
    const dim3 g(blocks(np, 128), nvec);
    if (fp32)
        hipLaunchKernelGGL((gamma_overlap_fix<float, float>), g, dim3(128), 0, st, np, nvec,
                           static_cast<const float*>(beta), ldb, psi, ldpsi, static_cast<float*>(P), ldp);
    else
        hipLaunchKernelGGL((gamma_overlap_fix<double, double>), g, dim3(128), 0, st, np, nvec,
                           static_cast<const double*>(beta), ldb, psi, ldpsi, static_cast<double*>(P), ldp);
    return hipGetLastError();
}

hipError_t launch_g0_real(int ncol, double* X, long ldx, hipStream_t st) {
    if (ncol <= 0) return hipSuccess;
    hipLaunchKernelGGL(g0_real, dim3(blocks(ncol, 128)), dim3(128), 0, st, ncol, X, ldx);
    return hipGetLastError();
}

hipError_t launch_gamma_norm_fix(int ncol, const double* X, long ldx, double* n2, hipStream_t st) {
    if (ncol <= 0) return hipSuccess;
    hipLaunchKernelGGL(gamma_norm_fix, dim3(blocks(ncol, 128)), dim3(128), 0, st, ncol, X, ldx, n2);
    return hipGetLastError();
}

hipError_t launch_zero_imag(int rows, int ncol, double* X, long ldx, hipStream_t st) {
    if (rows <= 0 || ncol <= 0) return hipSuccess;
    hipLaunchKernelGGL(zero_imag, dim3(blocks(rows, 128), ncol), dim3(128), 0, st, rows, ncol, X, ldx);
    return hipGetLastError();
}

}  // namespace q3x
