// Local part of H|psi> for one k-point: kinetic diagonal + FFT sandwich with V(r), as three
// fused kernels over a z-stick decomposition of the wavefunction sphere:
//
//   zsticks_inverse : sphere -> z-sticks (gather of psi), 1D backward FFT along i3,
//                     written plane-major  Z[band][i3][stick]
//   plane_sandwich  : per (band, i3) plane: scatter sticks into the occupied i1 columns,
//                     backward FFT along i2 (columns), then per i2 row backward FFT along
//                     i1, times V(r)/N, forward FFT along i1, keep the occupied columns,
//                     forward FFT along i2, write the sticks back
//   zsticks_forward : forward FFT along i3 per stick, then the sphere gather fused with
//                     the kinetic term: hpsi[g] = T[g] psi[g] + box[g]
//
// This is serial.hpp:450-478 (kinetic, memset, scatter, FFTW_BACKWARD, x V/N, FFTW_FORWARD,
// gather) restated sphere-aware: only the S occupied z-sticks and the occupied i1 columns
// of each plane are transformed, and the full box is never written to HBM.
//
// T = double (FP64) or float (mixed precision: psi and hpsi stay complex128 in HBM, the box
// and transforms are complex64).
#include <hip/hip_runtime.h>

#include "fft_lds.h"
#include "hlocal.h"

namespace q3x {

template <typename T>
__device__ __forceinline__ void load_twiddles(cx<T>* dst, const cx<T>* src, int n, int tid, int nthr) {
    for (int i = tid; i < n; i += nthr) dst[i] = src[i];
}

template <typename T>
__global__ void __launch_bounds__(256) zsticks_inverse(HLocalArgs a, const double* __restrict__ psi, long ldpsi,
                                                       cx<T>* __restrict__ zbuf) {
    extern __shared__ __align__(16) unsigned char smem_raw[];
    cx<T>* tw = reinterpret_cast<cx<T>*>(smem_raw);
    cx<T>* buf = tw + a.p3.n;
    const int tid = threadIdx.x, nthr = blockDim.x;
    const int n3 = a.p3.n;
    const int s0 = blockIdx.x * a.sticks_per_block;
    const int nsb = min(a.sticks_per_block, a.nst - s0);
    const int band = blockIdx.y;
    load_twiddles(tw, reinterpret_cast<const cx<T>*>(a.tw3), n3, tid, nthr);
    for (int i = tid; i < nsb * n3; i += nthr) buf[i] = cx<T>{T(0), T(0)};
    __syncthreads();
    const double* src = psi + band * ldpsi * 2;
    const int j0 = a.stick_off[s0], j1 = a.stick_off[s0 + nsb];
    const int base = s0 * n3;
    for (int j = j0 + tid; j < j1; j += nthr) {
        const int p = a.ent_pw[j];
        buf[a.ent_pos[j] - base] = cx<T>{T(src[2 * p]), T(src[2 * p + 1])};
    }
    __syncthreads();
    fft_lds<T, 2>(buf, a.p3, tw, +1, nsb, 1, n3, tid, nthr);
    cx<T>* out = zbuf + static_cast<long>(band) * n3 * a.nst;
    for (int i = tid; i < nsb * n3; i += nthr) {
        const int z = i / nsb, sl = i - z * nsb;
        out[static_cast<long>(z) * a.nst + s0 + sl] = buf[sl * n3 + z];
    }
}

template <typename T>
__global__ void __launch_bounds__(512) plane_sandwich(HLocalArgs a, const T* __restrict__ vplane,
                                                      cx<T>* __restrict__ zbuf) {
    extern __shared__ __align__(16) unsigned char smem_raw[];
    const int n1 = a.p1.n, n2 = a.p2.n, n3 = a.p3.n;
    const int ncol = a.ncol;
    cx<T>* tw1 = reinterpret_cast<cx<T>*>(smem_raw);
    cx<T>* tw2 = tw1 + n1;
    cx<T>* P = tw2 + n2;           // [n2][ncol]
    cx<T>* rows = P + n2 * ncol;   // [rows_per_pass][n1]
    const int tid = threadIdx.x, nthr = blockDim.x;
    const int z = blockIdx.x, band = blockIdx.y;
    load_twiddles(tw1, reinterpret_cast<const cx<T>*>(a.tw1), n1, tid, nthr);
    load_twiddles(tw2, reinterpret_cast<const cx<T>*>(a.tw2), n2, tid, nthr);
    for (int i = tid; i < n2 * ncol; i += nthr) P[i] = cx<T>{T(0), T(0)};
    __syncthreads();
    cx<T>* zp = zbuf + (static_cast<long>(band) * n3 + z) * a.nst;
    for (int s = tid; s < a.nst; s += nthr) P[a.stick_i2[s] * ncol + a.stick_col[s]] = zp[s];
    __syncthreads();
    // backward along i2 for every occupied column
    fft_lds<T, 4>(P, a.p2, tw2, +1, ncol, ncol, 1, tid, nthr);
    const T* vz = vplane + static_cast<long>(z) * n2 * n1;
    const int rp = a.rows_per_pass;
    for (int r0 = 0; r0 < n2; r0 += rp) {
        const int nr = min(rp, n2 - r0);
        for (int i = tid; i < nr * n1; i += nthr) rows[i] = cx<T>{T(0), T(0)};
        __syncthreads();
        for (int i = tid; i < nr * ncol; i += nthr) {
            const int rr = i / ncol, c = i - rr * ncol;
            rows[rr * n1 + a.col_i1[c]] = P[(r0 + rr) * ncol + c];
        }
        __syncthreads();
        fft_lds<T, 4>(rows, a.p1, tw1, +1, nr, 1, n1, tid, nthr);
        for (int i = tid; i < nr * n1; i += nthr) {
            const int rr = i / n1, x = i - rr * n1;
            rows[i] = cscale(rows[i], vz[(r0 + rr) * n1 + x]);
        }
        __syncthreads();
        fft_lds<T, 4>(rows, a.p1, tw1, -1, nr, 1, n1, tid, nthr);
        for (int i = tid; i < nr * ncol; i += nthr) {
            const int rr = i / ncol, c = i - rr * ncol;
            P[(r0 + rr) * ncol + c] = rows[rr * n1 + a.col_i1[c]];
        }
        __syncthreads();
    }
    fft_lds<T, 4>(P, a.p2, tw2, -1, ncol, ncol, 1, tid, nthr);
    for (int s = tid; s < a.nst; s += nthr) zp[s] = P[a.stick_i2[s] * ncol + a.stick_col[s]];
}

template <typename T>
__global__ void __launch_bounds__(256) zsticks_forward(HLocalArgs a, const cx<T>* __restrict__ zbuf,
                                                       const double* __restrict__ psi, long ldpsi,
                                                       double* __restrict__ hpsi, long ldh) {
    extern __shared__ __align__(16) unsigned char smem_raw[];
    cx<T>* tw = reinterpret_cast<cx<T>*>(smem_raw);
    cx<T>* buf = tw + a.p3.n;
    const int tid = threadIdx.x, nthr = blockDim.x;
    const int n3 = a.p3.n;
    const int s0 = blockIdx.x * a.sticks_per_block;
    const int nsb = min(a.sticks_per_block, a.nst - s0);
    const int band = blockIdx.y;
    load_twiddles(tw, reinterpret_cast<const cx<T>*>(a.tw3), n3, tid, nthr);
    const cx<T>* in = zbuf + static_cast<long>(band) * n3 * a.nst;
    for (int i = tid; i < nsb * n3; i += nthr) {
        const int z = i / nsb, sl = i - z * nsb;
        buf[sl * n3 + z] = in[static_cast<long>(z) * a.nst + s0 + sl];
    }
    __syncthreads();
    fft_lds<T, 2>(buf, a.p3, tw, -1, nsb, 1, n3, tid, nthr);
    const double* src = psi + band * ldpsi * 2;
    double* dst = hpsi + band * ldh * 2;
    const int j0 = a.stick_off[s0], j1 = a.stick_off[s0 + nsb];
    const int base = s0 * n3;
    for (int j = j0 + tid; j < j1; j += nthr) {
        const int p = a.ent_pw[j];
        const double t = a.ekin[p];
        const cx<T> v = buf[a.ent_pos[j] - base];
        dst[2 * p] = t * src[2 * p] + static_cast<double>(v.x);
        dst[2 * p + 1] = t * src[2 * p + 1] + static_cast<double>(v.y);
    }
}

// Plane waves that fall outside the wave grid (fft_map = -1) get the kinetic term only,
// as in serial.hpp:458-477 where the scatter and gather skip them.
__global__ void kinetic_only(const int* __restrict__ idx, int n, const double* __restrict__ ekin,
                             const double* __restrict__ psi, long ldpsi, double* __restrict__ hpsi, long ldh) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int p = idx[i];
    const long b = blockIdx.y;
    hpsi[2 * (b * ldh + p)] = ekin[p] * psi[2 * (b * ldpsi + p)];
    hpsi[2 * (b * ldh + p) + 1] = ekin[p] * psi[2 * (b * ldpsi + p) + 1];
}

template <typename T>
hipError_t launch_hlocal_t(const HLocalArgs& a, const T* vplane, void* zbuf, const double* psi, long ldpsi,
                           double* hpsi, long ldh, int nvec, hipStream_t st, hipEvent_t* marks) {
    if (marks) (void)hipEventRecord(marks[0], st);
    const size_t cs = sizeof(cx<T>);
    const int nsb_blocks = (a.nst + a.sticks_per_block - 1) / a.sticks_per_block;
    dim3 g1(nsb_blocks, nvec);
    const size_t lds_z = cs * (a.p3.n + static_cast<size_t>(a.sticks_per_block) * a.p3.n);
    hipLaunchKernelGGL(zsticks_inverse<T>, g1, dim3(256), lds_z, st, a, psi, ldpsi, static_cast<cx<T>*>(zbuf));
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    if (marks) (void)hipEventRecord(marks[1], st);
    const size_t lds_p = cs * (a.p1.n + a.p2.n + static_cast<size_t>(a.p2.n) * a.ncol +
                               static_cast<size_t>(a.rows_per_pass) * a.p1.n);
    static bool lds_raised = false;
    if (!lds_raised && lds_p > 65536) {
        e = hipFuncSetAttribute(reinterpret_cast<const void*>(&plane_sandwich<T>),
                                hipFuncAttributeMaxDynamicSharedMemorySize, static_cast<int>(lds_p));
        if (e != hipSuccess) return e;
        lds_raised = true;
    }
    hipLaunchKernelGGL(plane_sandwich<T>, dim3(a.p3.n, nvec), dim3(512), lds_p, st, a, vplane,
                       static_cast<cx<T>*>(zbuf));
    e = hipGetLastError();
    if (e != hipSuccess) return e;
    if (marks) (void)hipEventRecord(marks[2], st);
    hipLaunchKernelGGL(zsticks_forward<T>, g1, dim3(256), lds_z, st, a, static_cast<const cx<T>*>(zbuf), psi, ldpsi,
                       hpsi, ldh);
    e = hipGetLastError();
    if (e != hipSuccess) return e;
    if (marks) (void)hipEventRecord(marks[3], st);
    if (a.n_outside > 0) {
        hipLaunchKernelGGL(kinetic_only, dim3((a.n_outside + 255) / 256, nvec), dim3(256), 0, st, a.outside,
                           a.n_outside, a.ekin, psi, ldpsi, hpsi, ldh);
        e = hipGetLastError();
    }
    return e;
}

size_t hlocal_lds_bytes(const HLocalArgs& a, bool fp32, int which) {
    const size_t cs = fp32 ? 8 : 16;
    if (which == 0) return cs * (a.p3.n + static_cast<size_t>(a.sticks_per_block) * a.p3.n);
    return cs * (a.p1.n + a.p2.n + static_cast<size_t>(a.p2.n) * a.ncol + static_cast<size_t>(a.rows_per_pass) * a.p1.n);
}

hipError_t launch_hlocal(const HLocalArgs& a, bool fp32, const void* vplane, void* zbuf, const double* psi,
                         long ldpsi, double* hpsi, long ldh, int nvec, hipStream_t st, hipEvent_t* marks) {
    if (fp32)
        return launch_hlocal_t<float>(a, static_cast<const float*>(vplane), zbuf, psi, ldpsi, hpsi, ldh, nvec, st,
                                      marks);
    return launch_hlocal_t<double>(a, static_cast<const double*>(vplane), zbuf, psi, ldpsi, hpsi, ldh, nvec, st,
                                   marks);
}

}  // namespace q3x
