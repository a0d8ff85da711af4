// Complex GEMM on CDNA4 matrix cores (v_mfma_f64_16x16x4f64 / v_mfma_f32_16x16x4f32).
//
//   C (m x n) = alpha * op(A) * B + beta * C,   op(A) = A (m x k) or A^H (A stored k x m),
//   alpha, beta real, all matrices column-major interleaved complex.
//
// A complex product is run as one real product of doubled size: with A' (m x 2k) holding
// (Re, Im) of op(A) in adjacent k columns and B' (2k x 2n) built on the fly from B by the
// 2x2 block [[Br, Bi], [-Bi, Br]], A' B' = [Re C, Im C] interleaved -- 8mnk real flops, the
// same count as a complex GEMM.  Conjugation of op(A) = A^H is folded into the LDS load; B' is
// written to LDS (complex64) or derived from B at the fragment read (FP64).
//
// Workgroup tile: up to 128 rows x 80 complex columns, 4 wave64s in a 2x2 layout, each wave
// up to 4 x 5 MFMA 16x16 tiles; K staged through double-buffered LDS in chunks of 8 complex.
// Tall-skinny products (k >> m, n: projector overlaps, subspace Gram blocks) split
// K over the grid's z dimension into a workspace and are summed in a fixed order by a
// second kernel, so results do not depend on scheduling.
#include <hip/hip_runtime.h>

#include "zgemm.h"

#include <atomic>
#include <cstdio>
#include <cstdlib>
#include <type_traits>

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

constexpr int KC_DEFAULT = 8;  // complex k per LDS stage (16 real k, four MFMA k-steps)

// Workgroup: 2 x 2 wave64s; each wave owns TMW x TNW MFMA tiles of 16 x 16 real, so the
// workgroup tile is BM = 32 TMW rows by BN = 32 TNW real columns (16 TNW complex columns).
// K runs in stages of KC complex: the next stage's global loads are issued into registers
// before the current stage's MFMAs, and stored into the other LDS buffer after them, so one
// barrier per stage orders everything.  LDS rows are padded so that the k-contiguous
// operand stores (threads walking k) and the fragment reads (lanes walking rows / columns)
// both spread over the banks.
template <typename TA, typename TB, typename TM, typename TC, int OPA, int TMW, int TNW, int KC, int MINW>
__global__ void __launch_bounds__(256, MINW) zgemm_kernel(int m, int n, int k, int kchunk, const TA* __restrict__ A,
                                                    long lda, const TB* __restrict__ B, long ldb, TC* __restrict__ C,
                                                    long ldc, double alpha, double beta, TM* __restrict__ work,
                                                    int xmap) {
    using acc_t = typename Mma<TM>::acc_t;
    constexpr int BM = 32 * TMW;                 // rows of C
    constexpr int BNC = 16 * TNW;                // complex columns of C
    constexpr int APAD = sizeof(TM) == 8 ? 2 : 4;
    constexpr int LA = BM + APAD;                // As row stride (TM elements)
    constexpr int LB = 2 * BNC + 1;              // Bs row stride (real elements of B')
    constexpr int NLA = (BM * KC + 255) / 256;   // A loads per thread per stage
    constexpr int NLB = (BNC * KC + 255) / 256;  // B loads per thread per stage
    __shared__ TM As[2][2 * KC * LA];
    // FP64 (BC): B stored once per k, row p = [Br Bi Br Bi ...]; row 2p of B' is row p and row
    // 2p + 1 is row p with each (Br, Bi) pair read swapped and its real part negated.  Half the
    // LDS of B' (the 2x5 tiles reach three workgroups per CU instead of two): cfg2 projector
    // overlap 0.373 -> 0.349 ms, update 0.360 -> 0.340 ms.  Complex64 keeps B' expanded (its
    // overlap ran 5% slower with the compact form)
    constexpr bool BC = sizeof(TM) == 8;
    __shared__ TM Bs[2][(BC ? 1 : 2) * KC * LB];
    const int tid = threadIdx.x;
    const int lane = tid & 63, wave = tid >> 6;
    const int wm = wave & 1, wn = wave >> 1;
    const int li = lane & 15, lk = lane >> 4;
    // xmap (split K, gridDim.z a multiple of 8): workgroups are dispatched to the 8 XCDs
    // round-robin by linear id, so linear id i runs on XCD i mod 8; the tiles of K chunk z all go
    // to XCD z mod 8, where they read the chunk's rows of A and B through one L2 (with the plain
    // order, tile t of every chunk ran on XCD t and each chunk was fetched by every XCD)
    int bx = blockIdx.x, by = blockIdx.y, bz = blockIdx.z;
    if (xmap == 1) {
        const int T = gridDim.x * gridDim.y;
        const int i = blockIdx.x + gridDim.x * (blockIdx.y + gridDim.y * blockIdx.z);
        const int r = i >> 3, t = r % T;
        bz = (i & 7) + 8 * (r / T);
        bx = t % gridDim.x;
        by = t / gridDim.x;
    } else if (xmap == 2) {
        // no split, gridDim.x a multiple of 8: the gridDim.y column tiles of row block x run one
        // after another on XCD x mod 8 and share its rows of A through that L2
        const int i = blockIdx.x + gridDim.x * blockIdx.y;
        const int r = i >> 3;
        by = r % gridDim.y;
        bx = (i & 7) + 8 * (r / gridDim.y);
        if (bx * BM >= m) return;  // padding of the row-block count to a multiple of 8
    }
    const int m0 = bx * BM, n0 = by * BNC;
    const int kbeg = bz * kchunk;
    const int kend = min(k, kbeg + kchunk);

    acc_t acc[TMW][TNW];
#pragma unroll
    for (int a = 0; a < TMW; ++a)
#pragma unroll
        for (int b = 0; b < TNW; ++b) acc[a][b] = acc_t{0, 0, 0, 0};

    TM ra[NLA][2], rb[NLB][2];
    auto load = [&](int k0) {
#pragma unroll
        for (int q = 0; q < NLA; ++q) {
            const int e = tid + 256 * q;
            int i, p;
            if (OPA == 0) {
                i = e % BM;
                p = e / BM;
            } else {
                p = e % KC;
                i = e / KC;
            }
            const int gi = m0 + i, gp = k0 + p;
            ra[q][0] = TM(0);
            ra[q][1] = TM(0);
            if (e < BM * KC && gi < m && gp < kend) {
                const long o = OPA == 0 ? 2 * (gi + gp * lda) : 2 * (gp + gi * lda);
                ra[q][0] = static_cast<TM>(A[o]);
                ra[q][1] = OPA == 0 ? static_cast<TM>(A[o + 1]) : -static_cast<TM>(A[o + 1]);
            }
        }
#pragma unroll
        for (int q = 0; q < NLB; ++q) {
            const int e = tid + 256 * q;
            const int p = e % KC, j = e / KC;
            const int gj = n0 + j, gp = k0 + p;
            rb[q][0] = TM(0);
            rb[q][1] = TM(0);
            if (e < BNC * KC && gj < n && gp < kend) {
                const long o = 2 * (gp + gj * ldb);
                rb[q][0] = static_cast<TM>(B[o]);
                rb[q][1] = static_cast<TM>(B[o + 1]);
            }
        }
    };
    auto store = [&](int buf) {
#pragma unroll
        for (int q = 0; q < NLA; ++q) {
            const int e = tid + 256 * q;
            if (e < BM * KC) {
                const int i = OPA == 0 ? e % BM : e / KC;
                const int p = OPA == 0 ? e / BM : e % KC;
                As[buf][(2 * p) * LA + i] = ra[q][0];
                As[buf][(2 * p + 1) * LA + i] = ra[q][1];
            }
        }
#pragma unroll
        for (int q = 0; q < NLB; ++q) {
            const int e = tid + 256 * q;
            if (e < BNC * KC) {
                const int p = e % KC, j = e / KC;
                if (BC) {
                    TM* b0 = &Bs[buf][p * LB + 2 * j];
                    b0[0] = rb[q][0];
                    b0[1] = rb[q][1];
                } else {
                    TM* b0 = &Bs[buf][(2 * p) * LB + 2 * j];
                    b0[0] = rb[q][0];
                    b0[1] = rb[q][1];
                    b0[LB] = -rb[q][1];
                    b0[LB + 1] = rb[q][0];
                }
            }
        }
    };

    int cur = 0;
    if (kbeg < kend) {
        load(kbeg);
        store(0);
    }
    __syncthreads();
    for (int k0 = kbeg; k0 < kend; k0 += KC) {
        const bool more = k0 + KC < kend;
        if (more) load(k0 + KC);
        const TM* as = As[cur];
        const TM* bs = Bs[cur];
#pragma unroll
        for (int kk = 0; kk < (2 * KC) / 4; ++kk) {
            const int kr = 4 * kk + lk;
            TM af[TMW], bf[TNW];
#pragma unroll
            for (int t = 0; t < TMW; ++t) af[t] = as[kr * LA + wm * (16 * TMW) + t * 16 + li];
#pragma unroll
            for (int t = 0; t < TNW; ++t) {
                if (BC) {
                    // B' row kr = row kr / 2 of B; odd rows read the other entry of the pair, the
                    // real part negated at even columns (lane li's column parity is li's)
                    const TM x = bs[(kr >> 1) * LB + ((wn * (16 * TNW) + t * 16 + li) ^ (lk & 1))];
                    bf[t] = (lk & 1) && !(li & 1) ? -x : x;
                } else {
                    bf[t] = bs[kr * LB + wn * (16 * TNW) + t * 16 + li];
                }
            }
#pragma unroll
            for (int a = 0; a < TMW; ++a)
#pragma unroll
                for (int b = 0; b < TNW; ++b) acc[a][b] = Mma<TM>::run(af[a], bf[b], acc[a][b]);
        }
        if (more) store(cur ^ 1);
        __syncthreads();
        cur ^= 1;
    }

    // ---- epilogue: lane holds D[row(lk, r)][li] of each 16x16 tile
#pragma unroll
    for (int a = 0; a < TMW; ++a)
#pragma unroll
        for (int b = 0; b < TNW; ++b) {
            const int jr = wn * (16 * TNW) + b * 16 + li;
            const int gj = n0 + (jr >> 1), part = jr & 1;
            if (gj >= n) continue;
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int gi = m0 + wm * (16 * TMW) + a * 16 + Mma<TM>::row(lk, r);
                if (gi >= m) continue;
                const TM v = acc[a][b][r];
                if (work) {
                    work[(static_cast<long>(bz) * n + gj) * 2 * m + 2 * gi + part] = v;
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
    // partials summed in order z = 0, 1, ... (deterministic); eight loads are issued ahead of
    // their adds (as one load, wait, add per partial the loop waited out every load)
    double s = 0.0;
    int z = 0;
    for (; z + 8 <= ksplit; z += 8) {
        TM v[8];
#pragma unroll
        for (int u = 0; u < 8; ++u) v[u] = work[static_cast<long>(z + u) * total + idx];
#pragma unroll
        for (int u = 0; u < 8; ++u) s += static_cast<double>(v[u]);
    }
    for (; z < ksplit; ++z) s += static_cast<double>(work[static_cast<long>(z) * total + idx]);
    const long o = 2L * (rem >> 1) + 2L * j * ldc + (rem & 1);
    double out = alpha * s;
    if (beta != 0.0) out += beta * static_cast<double>(C[o]);
    C[o] = static_cast<TC>(out);
}

template <typename TA, typename TB, typename TM, typename TC, int TMW, int TNW, int KC = KC_DEFAULT, int MINW = 1>
hipError_t zgemm_shape(int opa, int m, int n, int k, double alpha, const void* A, long lda, const void* B, long ldb,
                       double beta, void* C, long ldc, void* work, size_t work_bytes, hipStream_t st,
                       int force_split = 0) {
    constexpr int BM = 32 * TMW, BNC = 16 * TNW;
    const int tm = (m + BM - 1) / BM, tn = (n + BNC - 1) / BNC;
    const long tiles = static_cast<long>(tm) * tn;
    int ksplit = 1;
    const long by_ws = work ? static_cast<long>(work_bytes / (2ull * m * n * sizeof(TM))) : 1;
    if (force_split < 0 && work && k > 0) {
        // fill one round of the GPU: workgroups resident at once = CUs x blocks per CU of this kernel
        static std::atomic<int> slots_cached{0};  // per kernel instantiation; same value from any thread
        int slots = slots_cached.load();
        if (slots == 0) {
            int dev = 0, cus = 0, per_cu = 0;
            (void)hipGetDevice(&dev);
            (void)hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
            (void)hipOccupancyMaxActiveBlocksPerMultiprocessor(
                &per_cu, reinterpret_cast<const void*>(&zgemm_kernel<TA, TB, TM, TC, 1, TMW, TNW, KC, MINW>), 256, 0);
            slots = std::max(1, cus * std::max(1, per_cu));
            slots_cached.store(slots);
        }
        const long by_k = std::max(1, k / 128);  // at least 128 complex k per split
        ksplit = static_cast<int>(std::max(1L, std::min(std::min(slots / tiles, by_k), by_ws)));
    } else if (force_split > 0 && work) {
        ksplit = static_cast<int>(std::max(1L, std::min<long>(force_split, by_ws)));
    } else if (k > 0 && tiles < 512 && work) {
        long want = (1024 + tiles - 1) / tiles;
        long by_k = (k + 511) / 512;  // at least 512 complex k per split
        ksplit = static_cast<int>(std::max(1L, std::min(want, std::min(by_k, by_ws))));
    }
    // FP64: split K over a multiple of 8 chunks when there are 8 or more, so that each chunk's
    // tiles share one XCD (xmap in zgemm_kernel).  At cfg2 the projector overlap then reads 347 MB
    // (the algorithmic 343 MB) instead of 977 MB in the same time: the kernel is not bound by
    // HBM, and the traffic saved is left to concurrent work.  The complex64 products ran 4% slower
    // with it (1% with the chunk count kept a multiple of 8: 0.228 vs 0.225 ms for the cfg2
    // overlap) and keep the plain order.
    // The chunk count stays that multiple of 8 after the chunk length is rounded up to KC: chunks
    // past the end of K are empty and add zero partials.
    const bool want_xmap = sizeof(TM) == 8 && ksplit >= 8;
    const bool keep8 = want_xmap && force_split <= 0;
    if (keep8) ksplit -= ksplit % 8;
    int kchunk = (k + ksplit - 1) / ksplit;
    kchunk = ((kchunk + KC - 1) / KC) * KC;
    if (kchunk == 0) kchunk = KC;
    if (!keep8) ksplit = k > 0 ? (k + kchunk - 1) / kchunk : 1;
    int xmap = want_xmap && ksplit % 8 == 0 ? 1 : 0;
    dim3 grid(tm, tn, ksplit);
    // A B without a split (the projector update Hpsi += beta W): the column tiles of a row block
    // run one after another on one XCD (xmap 2), so each row block of A is fetched once: at cfg2
    // the FP64 update reads 392 MB instead of 1,270 MB (0.365 -> 0.360 ms), complex64 286 instead
    // of 802 MB in the same time
    if (opa == 0 && ksplit == 1 && tn > 1) {
        xmap = 2;
        grid.x = (tm + 7) / 8 * 8;
    }
    TM* w = ksplit > 1 ? static_cast<TM*>(work) : nullptr;
    if (opa == 0)
        hipLaunchKernelGGL((zgemm_kernel<TA, TB, TM, TC, 0, TMW, TNW, KC, MINW>), grid, dim3(256), 0, st, m, n, k, kchunk,
                           static_cast<const TA*>(A), lda, static_cast<const TB*>(B), ldb, static_cast<TC*>(C), ldc,
                           alpha, beta, w, xmap);
    else
        hipLaunchKernelGGL((zgemm_kernel<TA, TB, TM, TC, 1, TMW, TNW, KC, MINW>), grid, dim3(256), 0, st, m, n, k, kchunk,
                           static_cast<const TA*>(A), lda, static_cast<const TB*>(B), ldb, static_cast<TC*>(C), ldc,
                           alpha, beta, w, xmap);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess || ksplit == 1) return e;
    const long total = 2L * m * n;
    hipLaunchKernelGGL((splitk_reduce<TM, TC>), dim3(static_cast<unsigned>((total + 255) / 256)), dim3(256), 0, st, m,
                       n, ksplit, static_cast<const TM*>(work), static_cast<TC*>(C), ldc, alpha, beta);
    return hipGetLastError();
}

// Tile configurations (wave tiles TMW x TNW, K stage KC, minimum waves per SIMD) and how a
// product picks one.  The table comes from timing every candidate on the H|psi> and Davidson
// shapes of the 64-atom cell (q3_amd/tune_gemm.py).  Deep-k A^H B products split K so that
// the grid is one round of resident workgroups (CUs x occupancy of the kernel): a partial
// second round costs as much as a full one.  FP64 A^H B: 2x5 tiles when n fills 80-column
// tiles, else 2x2; FP64 A B (tall m): 4x2 tiles at two waves per SIMD; FP32: 2x2 tiles.
// Q3X_GEMM_TILE="TMWxTNW/ksplit" overrides the choice (ksplit -1: one round, 0: default).
enum class Tile { T2x2, T4x2, T2x5, T4x4, T4x5 };

template <typename TA, typename TB, typename TM, typename TC>
hipError_t zgemm_tile(Tile t, int split, int opa, int m, int n, int k, double alpha, const void* A, long lda,
                      const void* B, long ldb, double beta, void* C, long ldc, void* work, size_t work_bytes,
                      hipStream_t st) {
    switch (t) {
        case Tile::T2x2:
            return zgemm_shape<TA, TB, TM, TC, 2, 2, 8, 1>(opa, m, n, k, alpha, A, lda, B, ldb, beta, C, ldc, work,
                                                           work_bytes, st, split);
        case Tile::T4x2:
            return zgemm_shape<TA, TB, TM, TC, 4, 2, 8, 2>(opa, m, n, k, alpha, A, lda, B, ldb, beta, C, ldc, work,
                                                           work_bytes, st, split);
        case Tile::T2x5:
            return zgemm_shape<TA, TB, TM, TC, 2, 5, 8, 2>(opa, m, n, k, alpha, A, lda, B, ldb, beta, C, ldc, work,
                                                           work_bytes, st, split);
        case Tile::T4x4:
            return zgemm_shape<TA, TB, TM, TC, 4, 4, 8, 2>(opa, m, n, k, alpha, A, lda, B, ldb, beta, C, ldc, work,
                                                           work_bytes, st, split);
        case Tile::T4x5:
            return zgemm_shape<TA, TB, TM, TC, 4, 5, 8, 2>(opa, m, n, k, alpha, A, lda, B, ldb, beta, C, ldc, work,
                                                           work_bytes, st, split);
    }
    return hipErrorInvalidValue;
}

template <typename TA, typename TB, typename TM, typename TC>
hipError_t zgemm_t(int opa, int m, int n, int k, double alpha, const void* A, long lda, const void* B, long ldb,
                   double beta, void* C, long ldc, void* work, size_t work_bytes, hipStream_t st) {
    if (m <= 0 || n <= 0) return hipSuccess;
    if (const char* cfg = std::getenv("Q3X_GEMM_TILE")) {
        int tmw = 0, tnw = 0, sp = 0;
        if (std::sscanf(cfg, "%dx%d/%d", &tmw, &tnw, &sp) == 3) {  // sp: -1 one round, 0 default, > 0 fixed
            const Tile t = tmw == 4 ? (tnw == 2 ? Tile::T4x2 : tnw == 4 ? Tile::T4x4 : Tile::T4x5)
                                    : (tnw == 5 ? Tile::T2x5 : Tile::T2x2);
            return zgemm_tile<TA, TB, TM, TC>(t, sp, opa, m, n, k, alpha, A, lda, B, ldb, beta, C, ldc, work,
                                              work_bytes, st);
        }
    }
    const bool f64 = std::is_same<TM, double>::value;
    Tile t = Tile::T2x2;
    if (m > 64 && n > 32) {
        if (opa == 1)
            t = f64 && n % 80 == 0 ? Tile::T2x5 : Tile::T2x2;
        else
            t = f64 ? Tile::T4x2 : Tile::T2x2;
    }
    return zgemm_tile<TA, TB, TM, TC>(t, opa == 1 ? -1 : 0, opa, m, n, k, alpha, A, lda, B, ldb, beta, C, ldc, work,
                                      work_bytes, st);
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
