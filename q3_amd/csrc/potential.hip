// Restriction of a density-grid potential to the wavefunction grid in G space, on the device
// (compute_smooth_grid_potential_from_dense and inverse_fft_smooth_potential,
// hamiltonian_applicator/planewave_potential_utils.hpp:33-130, called by update_potential,
// serial.hpp:64-78, through compute_smooth_grid_potential, serial.hpp:852-872):
//
//   V(G)        = (1/N_dense) FFTW_FORWARD of V(r) on the dense box
//   Vs(G)       = V(G) for every G of the smooth box (wrapped index w -> Miller m = w when
//                 w <= n/2, else w - n: the +n/2 slot of an even box takes the +n/2 coefficient)
//   V_smooth(r) = Re FFTW_BACKWARD (unnormalised) of Vs on the smooth box
//
// Runs once per potential update, so the kernels favour simplicity: a full box is transformed
// axis by axis, each workgroup loading whole lines into LDS and running the Stockham engine of
// fft_lds.h on them.
#include <hip/hip_runtime.h>

#include <algorithm>

#include "fft_lds.h"
#include "potential.h"

namespace q3x {
namespace {

constexpr int kPad = lds_pad_shift<cx<double>>();

// Line `li` of axis `ax` of a C-order box (i1, i2, i3), i3 fastest: offset of its first point and
// the stride between its points.
__device__ __forceinline__ void line_of(int li, int n1, int n2, int n3, int ax, long& base, long& stride) {
    if (ax == 2) {
        base = static_cast<long>(li) * n3;
        stride = 1;
    } else if (ax == 1) {
        const int i1 = li / n3, i3 = li - i1 * n3;
        base = static_cast<long>(i1) * n2 * n3 + i3;
        stride = n3;
    } else {
        base = li;  // li = i2 * n3 + i3
        stride = static_cast<long>(n2) * n3;
    }
}

__global__ void __launch_bounds__(256) box_fft_lines(cx<double>* __restrict__ box, int n1, int n2, int n3, int ax,
                                                     FftPlan p, const cx<double>* __restrict__ tw, int sign, int lpb,
                                                     double scale) {
    extern __shared__ __align__(16) unsigned char smem_raw[];
    cx<double>* buf = reinterpret_cast<cx<double>*>(smem_raw);
    const int n = p.n, ls = padded_len(n, kPad);
    const int nlines = static_cast<int>(static_cast<long>(n1) * n2 * n3 / n);
    const int l0 = blockIdx.x * lpb;
    const int nl = min(lpb, nlines - l0);
    const int tid = threadIdx.x;
    for (int i = tid; i < nl * n; i += blockDim.x) {
        const int l = i / n, e = i - l * n;
        long base, stride;
        line_of(l0 + l, n1, n2, n3, ax, base, stride);
        buf[l * ls + pidx<kPad>(e)] = box[base + e * stride];
    }
    __syncthreads();
    fft_lds<double, 1, false, kPad>(buf, p, tw, sign, nl, 1, ls, tid, blockDim.x);
    for (int i = tid; i < nl * n; i += blockDim.x) {
        const int l = i / n, e = i - l * n;
        long base, stride;
        line_of(l0 + l, n1, n2, n3, ax, base, stride);
        const cx<double> v = buf[l * ls + pidx<kPad>(e)];
        box[base + e * stride] = cx<double>{v.x * scale, v.y * scale};
    }
}

__global__ void real_to_complex(const double* __restrict__ v, cx<double>* __restrict__ out, long n) {
    const long i = static_cast<long>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < n) out[i] = cx<double>{v[i], 0.0};
}

// smooth[w] = dense[wrap(m(w))] over the smooth box, m(w) = w <= n/2 ? w : w - n per axis
__global__ void gather_smooth(const cx<double>* __restrict__ dense, int d1, int d2, int d3, cx<double>* __restrict__ smooth,
                              int s1, int s2, int s3) {
    const long i = static_cast<long>(blockIdx.x) * blockDim.x + threadIdx.x;
    const long total = static_cast<long>(s1) * s2 * s3;
    if (i >= total) return;
    const int w3 = static_cast<int>(i % s3);
    const int w2 = static_cast<int>((i / s3) % s2);
    const int w1 = static_cast<int>(i / (static_cast<long>(s2) * s3));
    const int m1 = w1 <= s1 / 2 ? w1 : w1 - s1, m2 = w2 <= s2 / 2 ? w2 : w2 - s2, m3 = w3 <= s3 / 2 ? w3 : w3 - s3;
    const int x1 = m1 >= 0 ? m1 : m1 + d1, x2 = m2 >= 0 ? m2 : m2 + d2, x3 = m3 >= 0 ? m3 : m3 + d3;
    smooth[i] = dense[(static_cast<long>(x1) * d2 + x2) * d3 + x3];
}

__global__ void real_part(const cx<double>* __restrict__ in, double* __restrict__ out, long n) {
    const long i = static_cast<long>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < n) out[i] = in[i].x;
}

hipError_t box_fft(cx<double>* box, const int dims[3], const FftPlan* plans, const cx<double>* const* tw, int sign,
                   double scale, hipStream_t st) {
    for (int ax = 0; ax < 3; ++ax) {
        const int n = dims[ax];
        const long nlines = static_cast<long>(dims[0]) * dims[1] * dims[2] / n;
        const int ls = padded_len(n, kPad);
        const int lpb = std::max(1, std::min(16, 48 * 1024 / (ls * 16)));
        const size_t lds = static_cast<size_t>(lpb) * ls * sizeof(cx<double>);
        const unsigned blocks = static_cast<unsigned>((nlines + lpb - 1) / lpb);
        hipLaunchKernelGGL(box_fft_lines, dim3(blocks), dim3(256), lds, st, box, dims[0], dims[1], dims[2], ax,
                           plans[ax], tw[ax], sign, lpb, ax == 2 ? scale : 1.0);
        const hipError_t e = hipGetLastError();
        if (e != hipSuccess) return e;
    }
    return hipSuccess;
}

}  // namespace

hipError_t launch_restrict_potential(const double* v_dense, const int dense_dims[3], const FftPlan dense_plans[3],
                                     const void* const dense_tw[3], const int smooth_dims[3],
                                     const FftPlan smooth_plans[3], const void* const smooth_tw[3], void* work_dense,
                                     void* work_smooth, double* v_smooth, hipStream_t st) {
    const long nd = static_cast<long>(dense_dims[0]) * dense_dims[1] * dense_dims[2];
    const long ns = static_cast<long>(smooth_dims[0]) * smooth_dims[1] * smooth_dims[2];
    auto* D = static_cast<cx<double>*>(work_dense);
    auto* S = static_cast<cx<double>*>(work_smooth);
    hipLaunchKernelGGL(real_to_complex, dim3(static_cast<unsigned>((nd + 255) / 256)), dim3(256), 0, st, v_dense, D,
                       nd);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    const cx<double>* tdw[3] = {static_cast<const cx<double>*>(dense_tw[0]), static_cast<const cx<double>*>(dense_tw[1]),
                                static_cast<const cx<double>*>(dense_tw[2])};
    if ((e = box_fft(D, dense_dims, dense_plans, tdw, -1, 1.0 / static_cast<double>(nd), st)) != hipSuccess) return e;
    hipLaunchKernelGGL(gather_smooth, dim3(static_cast<unsigned>((ns + 255) / 256)), dim3(256), 0, st, D, dense_dims[0],
                       dense_dims[1], dense_dims[2], S, smooth_dims[0], smooth_dims[1], smooth_dims[2]);
    if ((e = hipGetLastError()) != hipSuccess) return e;
    const cx<double>* tsw[3] = {static_cast<const cx<double>*>(smooth_tw[0]),
                                static_cast<const cx<double>*>(smooth_tw[1]),
                                static_cast<const cx<double>*>(smooth_tw[2])};
    if ((e = box_fft(S, smooth_dims, smooth_plans, tsw, +1, 1.0, st)) != hipSuccess) return e;
    hipLaunchKernelGGL(real_part, dim3(static_cast<unsigned>((ns + 255) / 256)), dim3(256), 0, st, S, v_smooth, ns);
    return hipGetLastError();
}

}  // namespace q3x
