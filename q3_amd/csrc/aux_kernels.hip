// Small fused kernels around the H-apply and the Davidson iteration.
#include <hip/hip_runtime.h>

#include "aux_kernels.h"

namespace q3x {

// W = D P for block-diagonal D (one real p_a x p_a block per atom, column-major, packed in
// atom order) -- serial.hpp:587-679.  P, W: nproj x nvec complex.
template <typename T>
__global__ void dblock_apply(int nproj, int nvec, const int* __restrict__ pstart, const int* __restrict__ psize,
                             const long* __restrict__ poff, const double* __restrict__ D, const T* __restrict__ P,
                             T* __restrict__ W) {
    const long idx = static_cast<long>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (idx >= static_cast<long>(nproj) * nvec) return;
    const int v = static_cast<int>(idx / nproj);
    const int p = static_cast<int>(idx - static_cast<long>(v) * nproj);
    const int s = pstart[p], bs = psize[p];
    const double* blk = D + poff[p];
    const int row = p - s;
    double re = 0.0, im = 0.0;
    const T* col = P + 2L * (static_cast<long>(v) * nproj + s);
    for (int q = 0; q < bs; ++q) {
        const double d = blk[row + q * bs];
        re += d * static_cast<double>(col[2 * q]);
        im += d * static_cast<double>(col[2 * q + 1]);
    }
    W[2 * idx] = static_cast<T>(re);
    W[2 * idx + 1] = static_cast<T>(im);
}

hipError_t launch_dblock_apply(bool fp32, int nproj, int nvec, const int* pstart, const int* psize, const long* poff,
                               const double* D, const void* P, void* W, hipStream_t st) {
    const long total = static_cast<long>(nproj) * nvec;
    if (total == 0) return hipSuccess;
    dim3 g(static_cast<unsigned>((total + 255) / 256));
    if (fp32)
        hipLaunchKernelGGL(dblock_apply<float>, g, dim3(256), 0, st, nproj, nvec, pstart, psize, poff, D,
                           static_cast<const float*>(P), static_cast<float*>(W));
    else
        hipLaunchKernelGGL(dblock_apply<double>, g, dim3(256), 0, st, nproj, nvec, pstart, psize, poff, D,
                           static_cast<const double*>(P), static_cast<double*>(W));
    return hipGetLastError();
}

// Teter-style preconditioner of the correction columns (serial.hpp:1138-1156):
//   r_g /= max(0.5 (1 + x + sqrt(1 + (x-1)^2)), 1e-4),  x = T_g - e_c
__global__ void teter_precondition(int npw, int ncol, const double* __restrict__ ekin, const double* __restrict__ eps,
                                   double* __restrict__ R, long ldr) {
    const int g = blockIdx.x * blockDim.x + threadIdx.x;
    const int c = blockIdx.y;
    if (g >= npw || c >= ncol) return;
    const double x = ekin[g] - eps[c];
    double den = 0.5 * (1.0 + x + sqrt(1.0 + (x - 1.0) * (x - 1.0)));
    if (den < 1e-4) den = 1e-4;
    double* col = R + 2L * c * ldr;
    col[2 * g] /= den;
    col[2 * g + 1] /= den;
}

hipError_t launch_teter(int npw, int ncol, const double* ekin, const double* eps, double* R, long ldr,
                        hipStream_t st) {
    if (npw == 0 || ncol == 0) return hipSuccess;
    hipLaunchKernelGGL(teter_precondition, dim3((npw + 255) / 256, ncol), dim3(256), 0, st, npw, ncol, ekin, eps, R,
                       ldr);
    return hipGetLastError();
}

// Corrections when the Ritz vectors are columns of the basis (right after a collapse the
// subspace eigenvectors are the identity): R_c = P_c (HV[:, src_c] - e_c V[:, src_c]) with the
// Teter preconditioner P_c of teter_precondition, in one pass; the two GEMMs against X = I of
// the general case (serial.hpp:995-1132) reduce to this gather.
__global__ void residual_columns_teter(int npw, const int* __restrict__ src, const double* __restrict__ eps,
                                       const double* __restrict__ ekin, const double2* __restrict__ V, long ldv,
                                       const double2* __restrict__ HV, long ldh, double2* __restrict__ R,
                                       long ldr) {
    const int g = blockIdx.x * blockDim.x + threadIdx.x;
    const int c = blockIdx.y;
    if (g >= npw) return;
    const long b = src[c];
    const double e = eps[c];
    const double2 v = V[g + b * ldv], h = HV[g + b * ldh];
    const double x = ekin[g] - e;
    double den = 0.5 * (1.0 + x + sqrt(1.0 + (x - 1.0) * (x - 1.0)));
    if (den < 1e-4) den = 1e-4;
    R[g + c * ldr] = double2{(h.x - e * v.x) / den, (h.y - e * v.y) / den};
}

hipError_t launch_residual_columns_teter(int npw, int ncol, const int* src, const double* eps, const double* ekin,
                                         const double* V, long ldv, const double* HV, long ldh, double* R, long ldr,
                                         hipStream_t st) {
    if (npw == 0 || ncol == 0) return hipSuccess;
    hipLaunchKernelGGL(residual_columns_teter, dim3((npw + 255) / 256, ncol), dim3(256), 0, st, npw, src, eps, ekin,
                       reinterpret_cast<const double2*>(V), ldv, reinterpret_cast<const double2*>(HV), ldh,
                       reinterpret_cast<double2*>(R), ldr);
    return hipGetLastError();
}

// Residual columns and their squared norms in one pass:
//   R_c = HV x_c - e_c V x_c is formed by the caller as two GEMMs; this reduces |R_c|^2
//   deterministically (one workgroup per column, fixed-order tree).
__global__ void __launch_bounds__(256) column_norm2(int npw, const double* __restrict__ X, long ldx,
                                                    double* __restrict__ out) {
    __shared__ double red[256];
    const int c = blockIdx.x;
    const double* col = X + 2L * c * ldx;
    double s = 0.0;
    for (int g = threadIdx.x; g < npw; g += 256) s += col[2 * g] * col[2 * g] + col[2 * g + 1] * col[2 * g + 1];
    red[threadIdx.x] = s;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if (threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) out[c] = red[0];
}

hipError_t launch_column_norm2(int npw, int ncol, const double* X, long ldx, double* out, hipStream_t st) {
    if (ncol == 0) return hipSuccess;
    hipLaunchKernelGGL(column_norm2, dim3(ncol), dim3(256), 0, st, npw, X, ldx, out);
    return hipGetLastError();
}

// Y (npw x ncol) = X, with column c scaled by s[c] (real).  ld in complex elements.
__global__ void scale_copy_columns(int npw, const double* __restrict__ X, long ldx, const double* __restrict__ s,
                                   double* __restrict__ Y, long ldy) {
    const int g = blockIdx.x * blockDim.x + threadIdx.x;
    const int c = blockIdx.y;
    if (g >= npw) return;
    const double f = s ? s[c] : 1.0;
    Y[2 * (g + c * ldy)] = f * X[2 * (g + c * ldx)];
    Y[2 * (g + c * ldy) + 1] = f * X[2 * (g + c * ldx) + 1];
}

hipError_t launch_scale_copy(int npw, int ncol, const double* X, long ldx, const double* s, double* Y, long ldy,
                             hipStream_t st) {
    if (npw == 0 || ncol == 0) return hipSuccess;
    hipLaunchKernelGGL(scale_copy_columns, dim3((npw + 255) / 256, ncol), dim3(256), 0, st, npw, X, ldx, s, Y, ldy);
    return hipGetLastError();
}

// Y[:, c] = s[c] X[:, src[c]] (s null -> 1): picks and scales the subspace eigenvectors of the
// unconverged bands (serial.hpp:1020-1060).
__global__ void gather_columns(int rows, const int* __restrict__ src, const double* __restrict__ s,
                               const double* __restrict__ X, long ldx, double* __restrict__ Y, long ldy) {
    const int g = blockIdx.x * blockDim.x + threadIdx.x;
    const int c = blockIdx.y;
    if (g >= rows) return;
    const double f = s ? s[c] : 1.0;
    const long ox = 2 * (g + src[c] * ldx), oy = 2 * (g + c * ldy);
    Y[oy] = f * X[ox];
    Y[oy + 1] = f * X[ox + 1];
}

hipError_t launch_gather_columns(int rows, int ncol, const int* src, const double* s, const double* X, long ldx,
                                 double* Y, long ldy, hipStream_t st) {
    if (rows == 0 || ncol == 0) return hipSuccess;
    hipLaunchKernelGGL(gather_columns, dim3((rows + 255) / 256, ncol), dim3(256), 0, st, rows, src, s, X, ldx, Y,
                       ldy);
    return hipGetLastError();
}

// Y = a_c X + beta Y column by column (a_c = a[c], or 1 when a is null).  X may alias Y.
__global__ void axpby_columns(int npw, const double* __restrict__ a, const double* X, long ldx, double beta,
                              double* Y, long ldy) {
    const int g = blockIdx.x * blockDim.x + threadIdx.x;
    const int c = blockIdx.y;
    if (g >= npw) return;
    const double f = a ? a[c] : 1.0;
    const long ox = 2 * (g + c * ldx), oy = 2 * (g + c * ldy);
    Y[oy] = f * X[ox] + beta * Y[oy];
    Y[oy + 1] = f * X[ox + 1] + beta * Y[oy + 1];
}

hipError_t launch_axpby(int npw, int ncol, const double* a, const double* X, long ldx, double beta, double* Y,
                        long ldy, hipStream_t st) {
    if (npw == 0 || ncol == 0) return hipSuccess;
    hipLaunchKernelGGL(axpby_columns, dim3((npw + 255) / 256, ncol), dim3(256), 0, st, npw, a, X, ldx, beta, Y, ldy);
    return hipGetLastError();
}

// Plain streaming copy (16-byte vectors): the reference point for HBM byte counters.
__global__ void copy16(const double2* __restrict__ src, double2* __restrict__ dst, long n) {
    for (long i = static_cast<long>(blockIdx.x) * blockDim.x + threadIdx.x; i < n;
         i += static_cast<long>(gridDim.x) * blockDim.x)
        dst[i] = src[i];
}

hipError_t launch_copy16(const void* src, void* dst, long n16, hipStream_t st) {
    if (n16 <= 0) return hipSuccess;
    hipLaunchKernelGGL(copy16, dim3(4096), dim3(256), 0, st, static_cast<const double2*>(src),
                       static_cast<double2*>(dst), n16);
    return hipGetLastError();
}

// Zero-fill of a device range with a kernel (keeps memsets out of captured graphs).
__global__ void fill_zero(double* __restrict__ p, long n) {
    const long i = static_cast<long>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0.0;
}

hipError_t launch_zero(double* p, long ndouble, hipStream_t st) {
    if (ndouble <= 0) return hipSuccess;
    hipLaunchKernelGGL(fill_zero, dim3(static_cast<unsigned>((ndouble + 255) / 256)), dim3(256), 0, st, p, ndouble);
    return hipGetLastError();
}

// Unit vectors e_{idx[c]} in the columns of V (kinetic seed, davidson.hpp:1079-1090).
__global__ void set_unit_columns(const int* __restrict__ idx, int ncol, double* __restrict__ V, long ldv) {
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= ncol) return;
    V[2 * (idx[c] + static_cast<long>(c) * ldv)] = 1.0;
}

// n x n complex identity in X (leading dimension ldx).
__global__ void set_identity(int n, double* __restrict__ X, long ldx) {
    const long e = static_cast<long>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (e >= static_cast<long>(n) * n) return;
    const int i = static_cast<int>(e % n), j = static_cast<int>(e / n);
    X[2 * (i + j * ldx)] = i == j ? 1.0 : 0.0;
    X[2 * (i + j * ldx) + 1] = 0.0;
}

hipError_t launch_identity(int n, double* X, long ldx, hipStream_t st) {
    if (n <= 0) return hipSuccess;
    const long total = static_cast<long>(n) * n;
    hipLaunchKernelGGL(set_identity, dim3(static_cast<unsigned>((total + 255) / 256)), dim3(256), 0, st, n, X, ldx);
    return hipGetLastError();
}

hipError_t launch_unit_columns(const int* idx, int ncol, double* V, long ldv, hipStream_t st) {
    if (ncol == 0) return hipSuccess;
    hipLaunchKernelGGL(set_unit_columns, dim3((ncol + 63) / 64), dim3(64), 0, st, idx, ncol, V, ldv);
    return hipGetLastError();
}

}  // namespace q3x
