// Complex GEMM on CDNA4 matrix cores (v_mfma_f64_16x16x4f64 / v_mfma_f32_16x16x4f32).
//
//   C (m x n) = alpha * op(A) * B + beta * C,   op(A) = A (m x k) or A^H (A stored k x m),
//   alpha, beta real, all matrices column-major interleaved complex.
//
// A complex product is run as one real product of doubled size: with A' (m x 2k) holding
// (Re, Im) of op(A) in adjacent k columns and B' (2k x 2n) built on the fly from B by the
// 2x2 block [[Br, Bi], [-Bi, Br]], A' B' = [Re C, Im C] interleaved -- 8mnk real flops, the
// same count as a complex GEMM.  Conjugation of op(A) = A^H is folded into the LDS load.
//
// Workgroup tile: 64 rows x 32 complex columns (64 real), 4 wave64s in a 2x2 layout, each
// wave a 32x32 real block as 2x2 MFMA 16x16 tiles; K staged through LDS in chunks of 16
// complex.  Tall-skinny products (k >> m, n: projector overlaps, subspace Gram blocks) split
// K over the grid's z dimension into a workspace and are summed in a fixed order by a
// second kernel, so results do not depend on scheduling.
#include <hip/hip_runtime.h>

#include "zgemm.h"

namespace q3x {

namespace {

template <typename T>
struct Mma;
template <>
struct Mma<double> {
    typedef double acc_t __attribute__((ext_vector_type(4)));
    // accumulator register r of lane (li + 16 lk) holds row 4 r + lk (rows interleaved)
    static __device__ __forceinline__ int row(int lk, int r) { return 4 * r + lk; }
    static __device__ __forceinline__ acc_t run(double a, double b, acc_t c) {
        return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
    }
};
template <>
struct Mma<float> {
    typedef float acc_t __attribute__((ext_vector_type(4)));
    // accumulator register r of lane (li + 16 lk) holds row 4 lk + r (rows blocked)
    static __device__ __forceinline__ int row(int lk, int r) { return 4 * lk + r; }
    static __device__ __forceinline__ acc_t run(float a, float b, acc_t c) {
        return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
    }
};

constexpr int BM = 64;   // rows per workgroup
constexpr int BNC = 32;  // complex columns per workgroup
constexpr int KC = 16;   // complex k per LDS stage

template <typename TA, typename TB, typename TM, typename TC, int OPA>
__global__ void __launch_bounds__(256) zgemm_kernel(int m, int n, int k, int kchunk, const TA* __restrict__ A,
                                                    long lda, const TB* __restrict__ B, long ldb, TC* __restrict__ C,
                                                    long ldc, double alpha, double beta, TM* __restrict__ work) {
    using acc_t = typename Mma<TM>::acc_t;
    __shared__ TM As[2 * KC][BM];
    __shared__ TM Bs[KC][BNC][2];
    const int tid = threadIdx.x;
    const int lane = tid & 63, wave = tid >> 6;
    const int wm = wave & 1, wn = wave >> 1;
    const int li = lane & 15, lk = lane >> 4;
    const int m0 = blockIdx.x * BM, n0 = blockIdx.y * BNC;
    const int kbeg = blockIdx.z * kchunk;
    const int kend = min(k, kbeg + kchunk);

    acc_t acc[2][2];
#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
        for (int b = 0; b < 2; ++b) acc[a][b] = acc_t{0, 0, 0, 0};

    for (int k0 = kbeg; k0 < kend; k0 += KC) {
        // ---- stage op(A) tile: As[2p + part][i]
        if (OPA == 0) {
            for (int e = tid; e < BM * KC; e += 256) {
                const int i = e % BM, p = e / BM;
                const int gi = m0 + i, gp = k0 + p;
                TM re = 0, im = 0;
                if (gi < m && gp < kend) {
                    const long o = 2 * (gi + gp * lda);
                    re = static_cast<TM>(A[o]);
                    im = static_cast<TM>(A[o + 1]);
                }
                As[2 * p][i] = re;
                As[2 * p + 1][i] = im;
            }
        } else {
            for (int e = tid; e < BM * KC; e += 256) {
                const int p = e % KC, i = e / KC;
                const int gi = m0 + i, gp = k0 + p;
                TM re = 0, im = 0;
                if (gi < m && gp < kend) {
                    const long o = 2 * (gp + gi * lda);
                    re = static_cast<TM>(A[o]);
                    im = -static_cast<TM>(A[o + 1]);
                }
                As[2 * p][i] = re;
                As[2 * p + 1][i] = im;
            }
        }
        // ---- stage B tile: Bs[p][j] = (Re, Im)
        for (int e = tid; e < BNC * KC; e += 256) {
            const int p = e % KC, j = e / KC;
            const int gj = n0 + j, gp = k0 + p;
            TM re = 0, im = 0;
            if (gj < n && gp < kend) {
                const long o = 2 * (gp + gj * ldb);
                re = static_cast<TM>(B[o]);
                im = static_cast<TM>(B[o + 1]);
            }
            Bs[p][j][0] = re;
            Bs[p][j][1] = im;
        }
        __syncthreads();
#pragma unroll
        for (int kk = 0; kk < (2 * KC) / 4; ++kk) {
            const int kr = 4 * kk + lk;
            const int p = kr >> 1, kpar = kr & 1;
            TM af[2], bf[2];
#pragma unroll
            for (int t = 0; t < 2; ++t) af[t] = As[kr][wm * 32 + t * 16 + li];
#pragma unroll
            for (int t = 0; t < 2; ++t) {
                const int jr = wn * 32 + t * 16 + li;
                const int j = jr >> 1, cpar = jr & 1;
                const TM br = Bs[p][j][0], bi = Bs[p][j][1];
                // B' = [[Br, Bi], [-Bi, Br]] : row parity kpar, column parity cpar
                bf[t] = kpar == 0 ? (cpar == 0 ? br : bi) : (cpar == 0 ? -bi : br);
            }
#pragma unroll
            for (int a = 0; a < 2; ++a)
#pragma unroll
                for (int b = 0; b < 2; ++b) acc[a][b] = Mma<TM>::run(af[a], bf[b], acc[a][b]);
        }
        __syncthreads();
    }

    // ---- epilogue: lane holds D[row(lk, r)][li] of each 16x16 tile
#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
        for (int b = 0; b < 2; ++b) {
            const int jr = wn * 32 + b * 16 + li;
            const int gj = n0 + (jr >> 1), part = jr & 1;
            if (gj >= n) continue;
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int gi = m0 + wm * 32 + a * 16 + Mma<TM>::row(lk, r);
                if (gi >= m) continue;
                const TM v = acc[a][b][r];
                if (work) {
                    work[(static_cast<long>(blockIdx.z) * n + gj) * 2 * m + 2 * gi + part] = v;
                } else {
                    const long o = 2 * (gi + gj * ldc) + part;
                    double out = alpha * static_cast<double>(v);
                    if (beta != 0.0) out += beta * static_cast<double>(C[o]);
                    C[o] = static_cast<TC>(out);
                }
            }
        }
}

template <typename TM, typename TC>
__global__ void splitk_reduce(int m, int n, int ksplit, const TM* __restrict__ work, TC* __restrict__ C, long ldc,
                              double alpha, double beta) {
    const long idx = static_cast<long>(blockIdx.x) * blockDim.x + threadIdx.x;
    const long total = 2L * m * n;
    if (idx >= total) return;
    const int j = static_cast<int>(idx / (2L * m));
    const int rem = static_cast<int>(idx - static_cast<long>(j) * 2 * m);
    double s = 0.0;
    for (int z = 0; z < ksplit; ++z) s += static_cast<double>(work[static_cast<long>(z) * total + idx]);
    const long o = 2L * (rem >> 1) + 2L * j * ldc + (rem & 1);
    double out = alpha * s;
    if (beta != 0.0) out += beta * static_cast<double>(C[o]);
    C[o] = static_cast<TC>(out);
}

template <typename TA, typename TB, typename TM, typename TC>
hipError_t zgemm_t(int opa, int m, int n, int k, double alpha, const void* A, long lda, const void* B, long ldb,
                   double beta, void* C, long ldc, void* work, size_t work_bytes, hipStream_t st) {
    if (m <= 0 || n <= 0) return hipSuccess;
    const int tm = (m + BM - 1) / BM, tn = (n + BNC - 1) / BNC;
    const long tiles = static_cast<long>(tm) * tn;
    int ksplit = 1;
    if (k > 0 && tiles < 512 && work) {
        long want = (1024 + tiles - 1) / tiles;
        long by_k = (k + 511) / 512;  // at least 512 complex k per split
        long by_ws = static_cast<long>(work_bytes / (2ull * m * n * sizeof(TM)));
        ksplit = static_cast<int>(std::max(1L, std::min(want, std::min(by_k, by_ws))));
    }
    int kchunk = (k + ksplit - 1) / ksplit;
    kchunk = ((kchunk + KC - 1) / KC) * KC;
    if (kchunk == 0) kchunk = KC;
    ksplit = k > 0 ? (k + kchunk - 1) / kchunk : 1;
    dim3 grid(tm, tn, ksplit);
    TM* w = ksplit > 1 ? static_cast<TM*>(work) : nullptr;
    if (opa == 0)
        hipLaunchKernelGGL((zgemm_kernel<TA, TB, TM, TC, 0>), grid, dim3(256), 0, st, m, n, k, kchunk,
                           static_cast<const TA*>(A), lda, static_cast<const TB*>(B), ldb, static_cast<TC*>(C), ldc,
                           alpha, beta, w);
    else
        hipLaunchKernelGGL((zgemm_kernel<TA, TB, TM, TC, 1>), grid, dim3(256), 0, st, m, n, k, kchunk,
                           static_cast<const TA*>(A), lda, static_cast<const TB*>(B), ldb, static_cast<TC*>(C), ldc,
                           alpha, beta, w);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess || ksplit == 1) return e;
    const long total = 2L * m * n;
    hipLaunchKernelGGL((splitk_reduce<TM, TC>), dim3(static_cast<unsigned>((total + 255) / 256)), dim3(256), 0, st, m,
                       n, ksplit, static_cast<const TM*>(work), static_cast<TC*>(C), ldc, alpha, beta);
    return hipGetLastError();
}

}  // namespace

hipError_t zgemm(GemmTypes types, int opa, int m, int n, int k, double alpha, const void* A, long lda,
                 const void* B, long ldb, double beta, void* C, long ldc, void* work, size_t work_bytes,
                 hipStream_t st) {
    switch (types) {
        case GemmTypes::D_D_D_D:
            return zgemm_t<double, double, double, double>(opa, m, n, k, alpha, A, lda, B, ldb, beta, C, ldc, work,
                                                           work_bytes, st);
        case GemmTypes::F_D_F_F:
            return zgemm_t<float, double, float, float>(opa, m, n, k, alpha, A, lda, B, ldb, beta, C, ldc, work,
                                                        work_bytes, st);
        case GemmTypes::F_F_F_D:
            return zgemm_t<float, float, float, double>(opa, m, n, k, alpha, A, lda, B, ldb, beta, C, ldc, work,
                                                        work_bytes, st);
        case GemmTypes::D_D_F_D:
            return zgemm_t<double, double, float, double>(opa, m, n, k, alpha, A, lda, B, ldb, beta, C, ldc, work,
                                                          work_bytes, st);
        case GemmTypes::F_F_F_F:
            return zgemm_t<float, float, float, float>(opa, m, n, k, alpha, A, lda, B, ldb, beta, C, ldc, work,
                                                       work_bytes, st);
    }
    return hipErrorInvalidValue;
}

}  // namespace q3x
