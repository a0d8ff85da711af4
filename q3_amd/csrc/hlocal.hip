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
// Work split: HBM loads and stores are done by the whole workgroup (coalesced), the 1D
// transforms by single wave64s that own whole columns / rows / sticks, so a plane needs
// four workgroup barriers; the stages inside a transform only order the wave's own LDS
// accesses.
//
// T = double (FP64) or float (mixed precision: psi and hpsi stay complex128 in HBM, the box
// and transforms are complex64).
#include <hip/hip_runtime.h>

#include <atomic>

#include "fft_lds.h"
#include "hlocal.h"

namespace q3x {

template <typename T>
__device__ __forceinline__ void load_twiddles(cx<T>* dst, const cx<T>* src, int n, int tid, int nthr) {
    for (int i = tid; i < n; i += nthr) dst[i] = src[i];
}

__device__ __forceinline__ void wave_sync() {
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    __builtin_amdgcn_wave_barrier();
}

// Sticks of this block are split over its waves in contiguous runs; stick sl of the block
// sits at buf + sl * ls (padded storage).
template <typename T>
__device__ __forceinline__ void wave_stick_fft(cx<T>* buf, const FftPlan& p, const cx<T>* tw, int sign, int nsb,
                                               int ls, int wave, int nw, int lane) {
    const int per = (nsb + nw - 1) / nw;
    const int first = wave * per;
    const int cnt = min(per, nsb - first);
    if (cnt > 0) fft_lds<T, 1, true, lds_pad_shift<cx<T>>()>(buf + first * ls, p, tw, sign, cnt, 1, ls, lane, 64);
}

template <typename T>
__global__ void __launch_bounds__(256) zsticks_inverse(HLocalArgs a, const double* __restrict__ psi, long ldpsi,
                                                       cx<T>* __restrict__ zbuf) {
    extern __shared__ __align__(16) unsigned char smem_raw[];
    constexpr int PS = lds_pad_shift<cx<T>>();
    cx<T>* tw = reinterpret_cast<cx<T>*>(smem_raw);
    cx<T>* buf = tw + a.p3.n;
    const int tid = threadIdx.x, nthr = blockDim.x;
    const int lane = tid & 63, wave = tid >> 6, nw = nthr >> 6;
    const int n3 = a.p3.n, l3 = padded_len(n3, PS);
    const int s0 = blockIdx.x * a.sticks_per_block;
    const int nsb = min(a.sticks_per_block, a.nst - s0);
    const int band = blockIdx.y;
    load_twiddles(tw, reinterpret_cast<const cx<T>*>(a.tw3), n3, tid, nthr);
    for (int i = tid; i < nsb * l3; i += nthr) buf[i] = cx<T>{T(0), T(0)};
    __syncthreads();
    const double* src = psi + band * ldpsi * 2;
    const int j0 = a.stick_off[s0], j1 = a.stick_off[s0 + nsb];
    const int base = s0 * n3;
    for (int j = j0 + tid; j < j1; j += nthr) {
        const int p = a.ent_pw[j];
        const int pos = a.ent_pos[j] - base, sl = pos / n3, z = pos - sl * n3;
        buf[sl * l3 + pidx<PS>(z)] = cx<T>{T(src[2 * p]), T(src[2 * p + 1])};
    }
    __syncthreads();
    wave_stick_fft(buf, a.p3, tw, +1, nsb, l3, wave, nw, lane);
    __syncthreads();
    cx<T>* out = zbuf + static_cast<long>(band) * n3 * a.nst;
    for (int i = tid; i < nsb * n3; i += nthr) {
        const int z = i / nsb, sl = i - z * nsb;
        out[static_cast<long>(z) * a.nst + s0 + sl] = buf[sl * l3 + pidx<PS>(z)];
    }
}

template <typename T>
__global__ void __launch_bounds__(512) plane_sandwich(HLocalArgs a, const T* __restrict__ vplane,
                                                      cx<T>* __restrict__ zbuf) {
    extern __shared__ __align__(16) unsigned char smem_raw[];
    constexpr int PS = lds_pad_shift<cx<T>>();
    const int n1 = a.p1.n, n2 = a.p2.n, n3 = a.p3.n;
    const int l1 = padded_len(n1, PS), l2 = padded_len(n2, PS);
    const int ncol = a.ncol;
    const int tid = threadIdx.x, nthr = blockDim.x;
    const int lane = tid & 63, wave = tid >> 6, nw = nthr >> 6;
    const int rbw = a.rows_per_pass;  // rows per wave per pass
    cx<T>* tw1 = reinterpret_cast<cx<T>*>(smem_raw);
    cx<T>* tw2 = tw1 + n1;
    cx<T>* P = tw2 + n2;                             // [ncol][l2], columns contiguous (padded)
    cx<T>* rows = P + l2 * ncol + wave * rbw * l1;   // this wave's [rbw][l1]
    const int z = blockIdx.x, band = blockIdx.y;
    load_twiddles(tw1, reinterpret_cast<const cx<T>*>(a.tw1), n1, tid, nthr);
    load_twiddles(tw2, reinterpret_cast<const cx<T>*>(a.tw2), n2, tid, nthr);
    for (int i = tid; i < l2 * ncol; i += nthr) P[i] = cx<T>{T(0), T(0)};
    __syncthreads();
    cx<T>* zp = zbuf + (static_cast<long>(band) * n3 + z) * a.nst;
    for (int s = tid; s < a.nst; s += nthr) P[a.stick_col[s] * l2 + pidx<PS>(a.stick_i2[s])] = zp[s];
    __syncthreads();
    // backward along i2: wave w owns columns w, w + nw, ...
    const int my_cols = wave < ncol ? (ncol - 1 - wave) / nw + 1 : 0;
    if (my_cols > 0) fft_lds<T, 1, true, PS>(P + wave * l2, a.p2, tw2, +1, my_cols, 1, nw * l2, lane, 64);
    __syncthreads();
    // rows: wave w owns rows w, w + nw, ... in passes of rbw rows
    const T* vz = vplane + static_cast<long>(z) * n2 * n1;
    const int my_rows = wave < n2 ? (n2 - 1 - wave) / nw + 1 : 0;
    for (int k0 = 0; k0 < my_rows; k0 += rbw) {
        const int nr = min(rbw, my_rows - k0);
        for (int i = lane; i < nr * l1; i += 64) rows[i] = cx<T>{T(0), T(0)};
        wave_sync();
        for (int i = lane; i < nr * ncol; i += 64) {
            const int rr = i / ncol, c = i - rr * ncol;
            const int i2 = wave + nw * (k0 + rr);
            rows[rr * l1 + pidx<PS>(a.col_i1[c])] = P[c * l2 + pidx<PS>(i2)];
        }
        wave_sync();
        fft_lds<T, 1, true, PS>(rows, a.p1, tw1, +1, nr, 1, l1, lane, 64);
        for (int i = lane; i < nr * n1; i += 64) {
            const int rr = i / n1, x = i - rr * n1;
            const int i2 = wave + nw * (k0 + rr);
            cx<T>& e = rows[rr * l1 + pidx<PS>(x)];
            e = cscale(e, vz[i2 * n1 + x]);
        }
        wave_sync();
        fft_lds<T, 1, true, PS>(rows, a.p1, tw1, -1, nr, 1, l1, lane, 64);
        for (int i = lane; i < nr * ncol; i += 64) {
            const int rr = i / ncol, c = i - rr * ncol;
            const int i2 = wave + nw * (k0 + rr);
            P[c * l2 + pidx<PS>(i2)] = rows[rr * l1 + pidx<PS>(a.col_i1[c])];
        }
        wave_sync();
    }
    __syncthreads();
    if (my_cols > 0) fft_lds<T, 1, true, PS>(P + wave * l2, a.p2, tw2, -1, my_cols, 1, nw * l2, lane, 64);
    __syncthreads();
    for (int s = tid; s < a.nst; s += nthr) zp[s] = P[a.stick_col[s] * l2 + pidx<PS>(a.stick_i2[s])];
}

template <typename T>
__global__ void __launch_bounds__(256) zsticks_forward(HLocalArgs a, const cx<T>* __restrict__ zbuf,
                                                       const double* __restrict__ psi, long ldpsi,
                                                       double* __restrict__ hpsi, long ldh) {
    extern __shared__ __align__(16) unsigned char smem_raw[];
    constexpr int PS = lds_pad_shift<cx<T>>();
    cx<T>* tw = reinterpret_cast<cx<T>*>(smem_raw);
    cx<T>* buf = tw + a.p3.n;
    const int tid = threadIdx.x, nthr = blockDim.x;
    const int lane = tid & 63, wave = tid >> 6, nw = nthr >> 6;
    const int n3 = a.p3.n, l3 = padded_len(n3, PS);
    const int s0 = blockIdx.x * a.sticks_per_block;
    const int nsb = min(a.sticks_per_block, a.nst - s0);
    const int band = blockIdx.y;
    load_twiddles(tw, reinterpret_cast<const cx<T>*>(a.tw3), n3, tid, nthr);
    const cx<T>* in = zbuf + static_cast<long>(band) * n3 * a.nst;
    for (int i = tid; i < nsb * n3; i += nthr) {
        const int z = i / nsb, sl = i - z * nsb;
        buf[sl * l3 + pidx<PS>(z)] = in[static_cast<long>(z) * a.nst + s0 + sl];
    }
    __syncthreads();
    wave_stick_fft(buf, a.p3, tw, -1, nsb, l3, wave, nw, lane);
    __syncthreads();
    const double* src = psi + band * ldpsi * 2;
    double* dst = hpsi + band * ldh * 2;
    const int j0 = a.stick_off[s0], j1 = a.stick_off[s0 + nsb];
    const int base = s0 * n3;
    for (int j = j0 + tid; j < j1; j += nthr) {
        const int p = a.ent_pw[j];
        const double t = a.ekin[p];
        const int pos = a.ent_pos[j] - base, sl = pos / n3, zz = pos - sl * n3;
        const cx<T> v = buf[sl * l3 + pidx<PS>(zz)];
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

size_t hlocal_lds_bytes(const HLocalArgs& a, bool fp32, int which) {
    const size_t cs = fp32 ? 8 : 16;
    const int ps = fp32 ? lds_pad_shift<cx<float>>() : lds_pad_shift<cx<double>>();
    if (which == 0) return cs * (a.p3.n + static_cast<size_t>(a.sticks_per_block) * padded_len(a.p3.n, ps));
    // plane kernel: twiddles, padded plane, one padded row buffer per wave (8 waves)
    return cs * (a.p1.n + a.p2.n + static_cast<size_t>(padded_len(a.p2.n, ps)) * a.ncol +
                 static_cast<size_t>(8) * a.rows_per_pass * padded_len(a.p1.n, ps));
}

template <typename T>
hipError_t launch_hlocal_t(const HLocalArgs& a, const T* vplane, void* zbuf, const double* psi, long ldpsi,
                           double* hpsi, long ldh, int nvec, hipStream_t st, hipEvent_t* marks,
                           const PlaneTlConfig* tl, void* pglob, const StickTlConfig* tz) {
    if (marks) (void)hipEventRecord(marks[0], st);
    const bool fp32 = sizeof(T) == 4;
    const int nsb_blocks = (a.nst + a.sticks_per_block - 1) / a.sticks_per_block;
    dim3 g1(nsb_blocks, nvec);
    const size_t lds_z = hlocal_lds_bytes(a, fp32, 0);
    const size_t lds_p = hlocal_lds_bytes(a, fp32, 1);
    static std::atomic<bool> lds_raised{false};  // set once per process; the attribute call is idempotent
    hipError_t e;
    if (!lds_raised) {
        e = hipFuncSetAttribute(reinterpret_cast<const void*>(&plane_sandwich<T>),
                                hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024);
        if (e != hipSuccess) return e;
        e = hipFuncSetAttribute(reinterpret_cast<const void*>(&zsticks_inverse<T>),
                                hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024);
        if (e != hipSuccess) return e;
        e = hipFuncSetAttribute(reinterpret_cast<const void*>(&zsticks_forward<T>),
                                hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024);
        if (e != hipSuccess) return e;
        lds_raised = true;
    }
    const bool ztl = tz && tz->ok;
    if (ztl) {
        e = launch_sticks_tl<T>(a, +1, psi, ldpsi, static_cast<cx<T>*>(zbuf), nullptr, 0, *tz, nvec, st);
    } else {
        hipLaunchKernelGGL(zsticks_inverse<T>, g1, dim3(256), lds_z, st, a, psi, ldpsi, static_cast<cx<T>*>(zbuf));
        e = hipGetLastError();
    }
    if (e != hipSuccess) return e;
    if (marks) (void)hipEventRecord(marks[1], st);
    if (tl && tl->ok) {
        e = launch_plane_tl<T>(a, vplane, static_cast<cx<T>*>(zbuf), *tl, nvec, pglob, st);
    } else {
        hipLaunchKernelGGL(plane_sandwich<T>, dim3(a.p3.n, nvec), dim3(512), lds_p, st, a, vplane,
                           static_cast<cx<T>*>(zbuf));
        e = hipGetLastError();
    }
    if (e != hipSuccess) return e;
    if (marks) (void)hipEventRecord(marks[2], st);
    if (ztl) {
        e = launch_sticks_tl<T>(a, -1, psi, ldpsi, static_cast<cx<T>*>(zbuf), hpsi, ldh, *tz, nvec, st);
    } else {
        hipLaunchKernelGGL(zsticks_forward<T>, g1, dim3(256), lds_z, st, a, static_cast<const cx<T>*>(zbuf), psi,
                           ldpsi, hpsi, ldh);
        e = hipGetLastError();
    }
    if (e != hipSuccess) return e;
    if (marks) (void)hipEventRecord(marks[3], st);
    if (a.n_outside > 0) {
        hipLaunchKernelGGL(kinetic_only, dim3((a.n_outside + 255) / 256, nvec), dim3(256), 0, st, a.outside,
                           a.n_outside, a.ekin, psi, ldpsi, hpsi, ldh);
        e = hipGetLastError();
    }
    return e;
}

hipError_t launch_hlocal(const HLocalArgs& a, bool fp32, const void* vplane, void* zbuf, const double* psi,
                         long ldpsi, double* hpsi, long ldh, int nvec, hipStream_t st, hipEvent_t* marks,
                         const PlaneTlConfig* tl, void* pglob, const StickTlConfig* tz) {
    if (fp32)
        return launch_hlocal_t<float>(a, static_cast<const float*>(vplane), zbuf, psi, ldpsi, hpsi, ldh, nvec, st,
                                      marks, tl, pglob, tz);
    return launch_hlocal_t<double>(a, static_cast<const double*>(vplane), zbuf, psi, ldpsi, hpsi, ldh, nvec, st,
                                   marks, tl, pglob, tz);
}

}  // namespace q3x
