// Two-level register transforms for the local Hamiltonian (serial.hpp:456-478): the plane
// stage for boxes whose plane is square (n1 = n2 = N = A * B), and the z-stick stages for
// n3 = A * B, for the box sizes listed in kTlShapes.
//
// Per (band, i3) plane, as plane_sandwich in hlocal.hip: scatter the plane's z-stick values
// into the occupied i1 columns, backward FFT along i2, per i2 row backward FFT along i1,
// times V(r)/N, forward FFT along i1 keeping the occupied columns, forward FFT along i2,
// write the sticks back.  The difference is how a length-N transform runs: as two register
// stages (the "four-step" split N = A * B) instead of one LDS round trip per radix.
//
//   backward, natural -> digit-permuted:  n = B n1 + n2,  k = k1 + A k2
//     stage 1 (lane n2):  y[k1] = sum_n1 x[B n1 + n2] w_A^(n1 k1),   y[k1] *= w_N^(n2 k1)
//     stage 2 (lane k1):  X[k1 + A k2] = sum_n2 y_n2[k1] w_B^(n2 k2)
//   forward, digit-permuted -> natural:   n = n1 + A n2,  m = m2 + B m1
//     stage A (lane n1):  u[m2] = sum_n2 z[n1 + A n2] w_B^(-n2 m2),  u[m2] *= w_N^(-n1 m2)
//     stage B (lane m2):  Y[m2 + B m1] = sum_n1 u_n1[m2] w_A^(-n1 m1)
//
// Logical slot q of a transform is stored at q + q / B (one pad slot per block of B).  Stage 1
// and stage B touch slots {B j + lane}, stage 2 and stage A the block {B lane + j}, each lane
// writing back the slots it read, so every stage runs in place and only the hand-over between
// stages needs a barrier.  The element k = k1 + A k2 of a backward column sits at slot
// B k1 + k2, which is where the forward stage A of the same column expects it.  A row's
// backward stage 2, the V(r) multiply and the forward stage A use the same values of the same
// lane, so they run on registers with no LDS traffic between them.  Transform strides in LDS
// are odd, and lanes run across transforms fastest, so that the lanes of a wave hit distinct
// banks.
#include <hip/hip_runtime.h>

#include <atomic>
#include <limits>

#include "fft_lds.h"
#include "hlocal.h"

namespace q3x {

__device__ __forceinline__ void wave_order() {
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    __builtin_amdgcn_wave_barrier();
}

// PG = false: the plane P lives in LDS.  PG = true (planes too large for LDS): P lives in this
// workgroup's slot of pglob, [gridDim.y][gridDim.x][ncol][lc], which stays in L2 while the
// workgroup runs; the rows in flight stay in LDS.
//
// Work split: a wave owns groups of G = 64 / max(A, B) transforms (columns, then rows, then
// columns again) and carries each group through both stages with wave-level ordering only; the
// lanes of a group are contiguous in the wave (lane = g * R + index, R lanes per transform in a
// stage).  Waves run their groups independently, so one wave's LDS and ALU latency overlaps
// another's; block barriers separate only the scatter, the column pass, the row pass (a row
// needs every column), the forward column pass and the gather.  The first nrw waves have a
// buffer of G rows in flight each and share the row pass; every wave takes part in the column
// passes, so a workgroup can have more waves than row buffers fit beside the plane (the FP64
// plane at one workgroup per CU: the column passes then finish in one round of groups).
// Launch bound: 12 waves in complex64; 8 in FP64, whose stages with the prefetched stick values
// need more than the 168 registers that 12 waves allow.
constexpr int plane_tl_max_threads(bool fp64, int n) { return fp64 || n > 144 ? 512 : 768; }
// transforms per wave group: 64 / max(A, B), so that one group's stage-2 / stage-A lanes (A per
// transform) fit in a wave.  The FP64 column passes of planes of 72 and more take 64 / B
// columns instead, so that stage 1 / B (B lanes per transform) fills the wave, and run stage 2 /
// A in ceil(GC A / 64) sub-passes (cfg2, 96 = 12 x 8: 8 columns, two sub-passes; 0.98 vs 1.01 ms)
__host__ __device__ constexpr int plane_group_base(int a, int b) { return 64 / (a > b ? a : b); }
__host__ __device__ constexpr int plane_group_cols(bool fp64, int a, int b) {
    return fp64 && a * b >= 72 ? 64 / b : plane_group_base(a, b);
}
__host__ __device__ constexpr int plane_group_rows(int a, int b) { return plane_group_base(a, b); }
// Twiddles in plane_tl's LDS.  FP64: the stage 1 / B table [A][B] and the stage 2 / A table
// [B][SA] (SA = A rounded up to a multiple of 8), laid out as the lanes read them, so that a
// lane's reads of one table are one base address plus immediate offsets (w^(n2 k) read from one
// table of N needs an address per k): 0.98 -> 0.95 ms at cfg2, 190 instead of 218 VGPRs, the
// bank-conflict cycles about unchanged.  Complex64: one table of N (the split tables measured
// no faster there).
__host__ __device__ constexpr bool plane_tw_split(bool fp64) { return fp64; }
__host__ __device__ constexpr int plane_tw2_stride(int a) { return (a + 7) & ~7; }
__host__ __device__ constexpr int plane_tw_entries(bool fp64, int a, int b) {
    return plane_tw_split(fp64) ? a * b + b * plane_tw2_stride(a) : a * b;
}
// waves per SIMD the register allocation must allow: two complex64 workgroups of eight waves
// per CU need four (128 registers); the radix-15/16 complex64 shapes need more and get two
constexpr int plane_tl_min_waves(bool fp64, int a) { return fp64 || a > 12 ? 2 : 4; }

template <typename T, int A, int B, bool PG>
__global__ void __launch_bounds__(plane_tl_max_threads(sizeof(T) == 8, A * B),
                                  plane_tl_min_waves(sizeof(T) == 8, A)) plane_tl(HLocalArgs a, const T* __restrict__ vplane,
                                                 cx<T>* __restrict__ zbuf, const cx<T>* __restrict__ twg, int lc,
                                                 int lr, int nrw, int nplanes, cx<T>* __restrict__ pglob) {
    constexpr int N = A * B;
    constexpr int S = B + 1;  // stored length of a block of B slots
    constexpr int GC = plane_group_cols(sizeof(T) == 8, A, B), GR = plane_group_rows(A, B);
    constexpr int NSC = (GC * A + 63) / 64, NSR = (GR * A + 63) / 64;  // stage-2 / stage-A sub-passes
    extern __shared__ __align__(16) unsigned char smem_raw[];
    // twiddles w_N^j = exp(-2 pi i j / N).  Split (FP64): stage 1 / B (lane n2) reads w^(n2 k)
    // at tw1[k B + n2], stage 2 / A (task k1) reads w^(k1 m) at tw2[m SA + k1]; otherwise tw[j]
    constexpr bool split = plane_tw_split(sizeof(T) == 8);
    constexpr int SA = plane_tw2_stride(A);
    cx<T>* tw = reinterpret_cast<cx<T>*>(smem_raw);
    cx<T>* tw1 = tw;
    cx<T>* tw2 = tw + N;
    int* colmap = reinterpret_cast<int*>(tw + plane_tw_entries(sizeof(T) == 8, A, B));  // i1 -> column or -1
    cx<T>* lds_next = reinterpret_cast<cx<T>*>(colmap + ((N + 3) & ~3));
    const int ncol = a.ncol;
    cx<T>* P = PG ? pglob + static_cast<long>(blockIdx.x) * ncol * lc : lds_next;  // [ncol][lc]
    const int tid = threadIdx.x, nt = blockDim.x;
    const int lane = tid & 63, wave = tid >> 6, nw = nt >> 6;
    cx<T>* Rw = (PG ? lds_next : P + ncol * lc) + wave * GR * lr;  // this wave's [GR][lr] (waves < nrw)
    const cx<T> zero = {T(0), T(0)};
    // FP64 with A <= 12: stage 1 of every group of every plane runs on lane n2 = lane mod B, so
    // its twiddles w^(n2 k) are held in registers (44 VGPRs at A = 12, within the 2-wave budget;
    // A = 15, 16 would spill); 0.95 -> 0.93 ms at cfg2
    constexpr bool w1_regs = sizeof(T) == 8 && A <= 12;
    cx<T> w1r[A];
#pragma unroll
    for (int k = 0; k < A; ++k) w1r[k] = w1_regs ? twg[(lane % B) * k] : zero;
    auto w1 = [&](int n2, int k) -> cx<T> {
        if constexpr (w1_regs) return w1r[k];
        return split ? tw1[k * B + n2] : tw[n2 * k];
    };
    auto w2 = [&](int k1, int m) -> cx<T> { return split ? tw2[m * SA + k1] : tw[k1 * m]; };

    for (int i = tid; i < N; i += nt) {
        if (split) {
            const int k = i / B, n2 = i - k * B;
            tw1[i] = twg[n2 * k];
        } else {
            tw[i] = twg[i];
        }
        colmap[i] = -1;
    }
    if (split) {
        for (int i = tid; i < B * SA; i += nt) {
            const int m = i / SA, k1 = i - m * SA;
            if (k1 < A) tw2[i] = twg[k1 * m];
        }
    }
    __syncthreads();
    for (int c = tid; c < ncol; c += nt) colmap[a.col_i1[c]] = c;
    __syncthreads();
    // FP64 (persistent): LDS offset of the occupied column i1 = B j + (lane mod B) (or -1), which
    // row stage 1 reads and stage B writes for every row of every plane, held in registers
    int cofs[A];
    {
        const int ib = (tid & 63) % B;
#pragma unroll
        for (int j = 0; j < A; ++j) {
            const int c = colmap[B * j + ib];
            cofs[j] = c >= 0 ? c * lc : -1;
        }
    }
    auto col_offset = [&](int j, int ib) {
        if constexpr (sizeof(T) == 8) {
            return cofs[j];
        } else {
            const int c = colmap[B * j + ib];
            return c >= 0 ? c * lc : -1;
        }
    };

    // FP64 (one workgroup per CU): persistent over work items q = blockIdx.x + k gridDim.x (the
    // plane of each given by plane_of below); the stick values of the next plane are loaded into registers while the
    // current one is transformed, so their latency is hidden behind the passes (a block barrier
    // waits for LDS only), and the sticks' LDS slots stay in registers.  Complex64 (two
    // workgroups per CU, which overlap each other's loads): one plane per workgroup.
    constexpr bool persist = sizeof(T) == 8;
    constexpr int PF = persist ? 4 : 1;  // stick values per thread carried over; the rest load at once
    cx<T> pre[PF];  // (initialised: a partly written array would be kept in scratch memory)
#pragma unroll
    for (int i = 0; i < PF; ++i) pre[i] = zero;
    int soff[PF];  // LDS slot of stick tid + i nt, the same in every plane
#pragma unroll
    for (int i = 0; i < PF; ++i) {
        const int s = tid + i * nt;
        soff[i] = 0;
        if (persist && s < a.nst) {
            const int q = a.stick_i2[s];
            soff[i] = a.stick_col[s] * lc + q + q / B;
        }
    }
    // FP64: work items are spread over the XCDs by q mod 8 (workgroups are dispatched to the
    // eight XCDs round-robin and the grid is a multiple of 8), so when 8 divides n3, item q runs
    // band j mod nvec at i3 = 8 (j / nvec) + (q mod 8), j = q / 8: each XCD transforms only the
    // planes i3 = x mod 8, i3-major, its concurrent workgroups share one i3 and every V(r) plane
    // is read into one XCD's L2 once.  Otherwise band q mod nvec at i3 = q / nvec.  Both are
    // bijections of [0, nplanes).  Complex64: q = plane.
    const int nvec = nplanes / a.p3.n;
    const bool by_xcd = (a.p3.n & 7) == 0;
    auto plane_of = [&](int q) {
        if (!persist) return q;
        if (by_xcd) {
            const int j = q >> 3;
            return (j % nvec) * a.p3.n + (j / nvec) * 8 + (q & 7);
        }
        return (q % nvec) * a.p3.n + q / nvec;
    };
    auto prefetch = [&](int plane) {
        const cx<T>* src = zbuf + static_cast<long>(plane) * a.nst;
#pragma unroll
        for (int i = 0; i < PF; ++i) {
            const int s = tid + i * nt;
            if (s < a.nst) pre[i] = src[s];
        }
    };
    if (persist && blockIdx.x < nplanes) prefetch(plane_of(blockIdx.x));
    for (int q = blockIdx.x; q < nplanes; q += persist ? gridDim.x : nplanes) {
        const int plane = plane_of(q);
        const int z = plane % a.p3.n;
        for (int i = tid; i < ncol * lc; i += nt) P[i] = zero;
        __syncthreads();
        cx<T>* zp = zbuf + static_cast<long>(plane) * a.nst;
        if (persist) {
#pragma unroll
            for (int i = 0; i < PF; ++i)
                if (tid + i * nt < a.nst) P[soff[i]] = pre[i];
        }
        for (int s = tid + (persist ? PF * nt : 0); s < a.nst; s += nt) {
            const int q = a.stick_i2[s];
            P[a.stick_col[s] * lc + q + q / B] = zp[s];
        }
        __syncthreads();
        if (persist && q + static_cast<int>(gridDim.x) < nplanes) prefetch(plane_of(q + gridDim.x));

        // stage 1 / stage B: lane gB * B + n2 (B lanes per transform); stage 2 / stage A: task
        // t = lane + 64 s of sub-pass s is (transform t / A, index t mod A)
        const int gB = lane / B, iB = lane - gB * B;

        // columns, backward: stage 1 (lane n2 = iB), stage 2 (task k1)
        for (int c0 = wave * GC; c0 < ncol; c0 += nw * GC) {
            if (gB < GC && c0 + gB < ncol) {
                const int n2 = iB;
                cx<T>* col = P + (c0 + gB) * lc + n2;
                cx<T> v[A];
#pragma unroll
                for (int j = 0; j < A; ++j) v[j] = col[j * S];
                Dft<T, A>::run(v, T(1));
#pragma unroll
                for (int k = 1; k < A; ++k) v[k] = cmulc(v[k], w1(n2, k));
#pragma unroll
                for (int k = 0; k < A; ++k) col[k * S] = v[k];
            }
            wave_order();
#pragma unroll
            for (int sp = 0; sp < NSC; ++sp) {
                const int t = lane + 64 * sp, g = t / A;
                if (g < GC && c0 + g < ncol) {
                    cx<T>* blk = P + (c0 + g) * lc + (t - g * A) * S;
                    cx<T> u[B];
#pragma unroll
                    for (int j = 0; j < B; ++j) u[j] = blk[j];
                    Dft<T, B>::run(u, T(1));
#pragma unroll
                    for (int j = 0; j < B; ++j) blk[j] = u[j];
                }
            }
            wave_order();
        }
        __syncthreads();

        // rows; row i2 of the plane sits at slot (i2 % A) * S + i2 / A of each column
        for (int r0 = wave * GR; wave < nrw && r0 < N; r0 += nrw * GR) {
            // V(r) values of this lane's stage-2 tasks, loaded ahead of stage 1 so that their
            // latency overlaps it
            T vv[NSR][B];
#pragma unroll
            for (int sp = 0; sp < NSR; ++sp) {
                const int t = lane + 64 * sp, g = t / A, k1 = t - g * A;
                if (g < GR && r0 + g < N) {
                    const T* vrow = vplane + (static_cast<long>(z) * N + r0 + g) * N + k1;
#pragma unroll
                    for (int j = 0; j < B; ++j) vv[sp][j] = vrow[A * j];
                }
            }
            if (gB < GR && r0 + gB < N) {  // backward stage 1, lane n2
                const int n2 = iB, i2 = r0 + gB;
                const int pp = (i2 % A) * S + i2 / A;
                cx<T> v[A];
#pragma unroll
                for (int j = 0; j < A; ++j) {
                    const int o = col_offset(j, n2);
                    v[j] = zero;
                    if (o >= 0) v[j] = P[o + pp];
                }
                Dft<T, A>::run(v, T(1));
#pragma unroll
                for (int k = 1; k < A; ++k) v[k] = cmulc(v[k], w1(n2, k));
                cx<T>* row = Rw + gB * lr + n2;
#pragma unroll
                for (int k = 0; k < A; ++k) row[k * S] = v[k];
            }
            wave_order();
#pragma unroll
            for (int sp = 0; sp < NSR; ++sp) {  // backward stage 2, x V(r)/N, forward stage A; task (row, k1)
                const int t = lane + 64 * sp, g = t / A, k1 = t - g * A;
                if (g < GR && r0 + g < N) {
                    cx<T>* blk = Rw + g * lr + k1 * S;
                    cx<T> u[B];
#pragma unroll
                    for (int j = 0; j < B; ++j) u[j] = blk[j];
                    Dft<T, B>::run(u, T(1));
#pragma unroll
                    for (int j = 0; j < B; ++j) u[j] = cscale(u[j], vv[sp][j]);
                    Dft<T, B>::run(u, T(-1));
#pragma unroll
                    for (int m = 1; m < B; ++m) u[m] = cmul(u[m], w2(k1, m));
#pragma unroll
                    for (int j = 0; j < B; ++j) blk[j] = u[j];
                }
            }
            wave_order();
            if (gB < GR && r0 + gB < N) {  // forward stage B, lane m2; outputs i1 = m2 + B m1, occupied only
                const int m2 = iB, i2 = r0 + gB;
                const int pp = (i2 % A) * S + i2 / A;
                const cx<T>* row = Rw + gB * lr + m2;
                cx<T> v[A];
#pragma unroll
                for (int j = 0; j < A; ++j) v[j] = row[j * S];
                Dft<T, A>::run(v, T(-1));
#pragma unroll
                for (int m = 0; m < A; ++m) {
                    const int o = col_offset(m, m2);
                    if (o >= 0) P[o + pp] = v[m];
                }
            }
            wave_order();
        }
        __syncthreads();

        // columns, forward: stage A (task n1, block n1 * S holds elements n1 + A n2), stage B
        for (int c0 = wave * GC; c0 < ncol; c0 += nw * GC) {
#pragma unroll
            for (int sp = 0; sp < NSC; ++sp) {
                const int t = lane + 64 * sp, g = t / A;
                if (g < GC && c0 + g < ncol) {
                    const int n1 = t - g * A;
                    cx<T>* blk = P + (c0 + g) * lc + n1 * S;
                    cx<T> u[B];
#pragma unroll
                    for (int j = 0; j < B; ++j) u[j] = blk[j];
                    Dft<T, B>::run(u, T(-1));
#pragma unroll
                    for (int m = 1; m < B; ++m) u[m] = cmul(u[m], w2(n1, m));
#pragma unroll
                    for (int j = 0; j < B; ++j) blk[j] = u[j];
                }
            }
            wave_order();
            if (gB < GC && c0 + gB < ncol) {
                cx<T>* col = P + (c0 + gB) * lc + iB;
                cx<T> v[A];
#pragma unroll
                for (int j = 0; j < A; ++j) v[j] = col[j * S];
                Dft<T, A>::run(v, T(-1));
#pragma unroll
                for (int m = 0; m < A; ++m) col[m * S] = v[m];
            }
            wave_order();
        }
        __syncthreads();
        if (persist) {
#pragma unroll
            for (int i = 0; i < PF; ++i)
                if (tid + i * nt < a.nst) zp[tid + i * nt] = P[soff[i]];
        }
        for (int s = tid + (persist ? PF * nt : 0); s < a.nst; s += nt) {
            const int q = a.stick_i2[s];
            zp[s] = P[a.stick_col[s] * lc + q + q / B];
        }
        __syncthreads();  // P is cleared for the next plane
    }
}

// streaming (non-temporal) load of the stick buffer Z, which a pass reads once
template <typename T>
__device__ __forceinline__ cx<T> ld_stream(const cx<T>* p) {
    typedef T v2 __attribute__((ext_vector_type(2)));
    const v2 v = __builtin_nontemporal_load(reinterpret_cast<const v2*>(p));
    return cx<T>{v.x, v.y};
}

// Workgroup -> (stick block, band) for the z-stick kernels.  Workgroups are dispatched to the
// 8 XCDs round-robin by linear id, so id mod 8 names the XCD; giving every stick block of band b
// to XCD b mod 8 makes each band's psi / Hpsi column live in one L2, where the scattered 16-byte
// gathers and scatters of the sphere entries then hit (instead of every XCD fetching most of
// every band).  1D grid of 8 * nblk * ceil(nvec / 8) workgroups; the surplus ones return.
__device__ __forceinline__ bool stick_block(int nblk, int nvec, int& sb, int& band) {
    const int i = blockIdx.x;
    const int x = i & 7, r = i >> 3;
    const int q = r / nblk;
    band = x + 8 * q;
    sb = r - q * nblk;
    return band < nvec;
}

// z-stick transforms with the same two-level split, N = n3 = A * B.  The sticks of a block
// sit in LDS at sl * lz + slot(q), slot(q) = q + q / B, lanes running across sticks fastest.
//
// Inverse: scatter the sphere entries of psi, backward stage 1 in LDS, then backward stage 2
// in registers, written straight to Z[band][i3][stick] (element i3 = k1 + A k2 of lane k1).
template <typename T, int A, int B>
__global__ void __launch_bounds__(256) zsticks_inverse_tl(HLocalArgs a, const double* __restrict__ psi, long ldpsi,
                                                          cx<T>* __restrict__ zbuf, const cx<T>* __restrict__ twg,
                                                          int spb, int lz, int nblk, int nvec) {
    constexpr int N = A * B;
    constexpr int S = B + 1;
    extern __shared__ __align__(16) unsigned char smem_raw[];
    cx<T>* tw = reinterpret_cast<cx<T>*>(smem_raw);
    cx<T>* buf = tw + N;
    const int tid = threadIdx.x, nt = blockDim.x;
    int sb, band;
    if (!stick_block(nblk, nvec, sb, band)) return;
    const int s0 = sb * spb;
    const int nsb = min(spb, a.nst - s0);
    for (int i = tid; i < N; i += nt) tw[i] = twg[i];
    for (int i = tid; i < nsb * lz; i += nt) buf[i] = cx<T>{T(0), T(0)};
    __syncthreads();
    const double2* src = reinterpret_cast<const double2*>(psi + band * ldpsi * 2);
    const int j0 = a.stick_off[s0], j1 = a.stick_off[s0 + nsb];
    const int base = s0 * N;
    // two entries per thread per round: their index loads, then their scattered psi loads, are
    // issued together (cfg2 FP64 0.152 -> 0.145 ms; four or eight per round: 0.146, 0.148 ms;
    // the same prefetch of psi ahead of the forward kernel's transforms made it slower)
    constexpr int E = 2;
    for (int jb = j0 + tid; jb < j1; jb += E * nt) {
        int ep[E], eo[E];
#pragma unroll
        for (int e = 0; e < E; ++e) {
            const int j = jb + e * nt;
            ep[e] = -1;
            eo[e] = 0;
            if (j < j1) {
                ep[e] = a.ent_pw[j];
                const int pos = a.ent_pos[j] - base, sl = pos / N, z = pos - sl * N;
                eo[e] = sl * lz + z + z / B;
            }
        }
        double2 ex[E];
#pragma unroll
        for (int e = 0; e < E; ++e) ex[e] = ep[e] >= 0 ? src[ep[e]] : double2{0.0, 0.0};
#pragma unroll
        for (int e = 0; e < E; ++e)
            if (ep[e] >= 0) buf[eo[e]] = cx<T>{T(ex[e].x), T(ex[e].y)};
    }
    __syncthreads();
    for (int t = tid; t < nsb * B; t += nt) {
        const int n2 = t / nsb, sl = t - n2 * nsb;
        cx<T>* x = buf + sl * lz + n2;
        cx<T> v[A];
#pragma unroll
        for (int j = 0; j < A; ++j) v[j] = x[j * S];
        Dft<T, A>::run(v, T(1));
#pragma unroll
        for (int k = 1; k < A; ++k) v[k] = cmulc(v[k], tw[n2 * k]);
#pragma unroll
        for (int k = 0; k < A; ++k) x[k * S] = v[k];
    }
    __syncthreads();
    cx<T>* out = zbuf + static_cast<long>(band) * N * a.nst + s0;
    for (int t = tid; t < nsb * A; t += nt) {
        const int k1 = t / nsb, sl = t - k1 * nsb;
        const cx<T>* x = buf + sl * lz + k1 * S;
        cx<T> u[B];
#pragma unroll
        for (int j = 0; j < B; ++j) u[j] = x[j];
        Dft<T, B>::run(u, T(1));
#pragma unroll
        for (int j = 0; j < B; ++j) out[static_cast<long>(k1 + A * j) * a.nst + sl] = u[j];
    }
}

// Forward: stage A reads Z[band][n1 + A n2][stick] straight into registers (lane n1), stage B
// in LDS puts the transform in natural order, then the sphere gather fused with the kinetic
// term, hpsi[g] = T[g] psi[g] + box[g] (serial.hpp:451-478).
template <typename T, int A, int B>
__global__ void __launch_bounds__(256) zsticks_forward_tl(HLocalArgs a, const cx<T>* __restrict__ zbuf,
                                                          const double* __restrict__ psi, long ldpsi,
                                                          double* __restrict__ hpsi, long ldh,
                                                          const cx<T>* __restrict__ twg, int spb, int lz, int nblk,
                                                          int nvec) {
    constexpr int N = A * B;
    constexpr int S = B + 1;
    extern __shared__ __align__(16) unsigned char smem_raw[];
    cx<T>* tw = reinterpret_cast<cx<T>*>(smem_raw);
    cx<T>* buf = tw + N;
    const int tid = threadIdx.x, nt = blockDim.x;
    int sb, band;
    if (!stick_block(nblk, nvec, sb, band)) return;
    const int s0 = sb * spb;
    const int nsb = min(spb, a.nst - s0);
    for (int i = tid; i < N; i += nt) tw[i] = twg[i];
    __syncthreads();
    const cx<T>* in = zbuf + static_cast<long>(band) * N * a.nst + s0;
    for (int t = tid; t < nsb * A; t += nt) {
        const int n1 = t / nsb, sl = t - n1 * nsb;
        cx<T> u[B];
#pragma unroll
        for (int j = 0; j < B; ++j) {
            // FP64: Z read as a stream, so that it does not push this band's psi and Hpsi lines
            // (gathered and scattered below) out of L2 (cfg2 0.245 -> 0.226 ms; complex64 ran
            // 9% slower with it)
            const cx<T>* q = in + static_cast<long>(n1 + A * j) * a.nst + sl;
            u[j] = sizeof(T) == 8 ? ld_stream(q) : *q;
        }
        Dft<T, B>::run(u, T(-1));
#pragma unroll
        for (int m = 1; m < B; ++m) u[m] = cmul(u[m], tw[n1 * m]);
        cx<T>* x = buf + sl * lz + n1 * S;
#pragma unroll
        for (int m = 0; m < B; ++m) x[m] = u[m];
    }
    __syncthreads();
    for (int t = tid; t < nsb * B; t += nt) {
        const int m2 = t / nsb, sl = t - m2 * nsb;
        cx<T>* x = buf + sl * lz + m2;
        cx<T> v[A];
#pragma unroll
        for (int j = 0; j < A; ++j) v[j] = x[j * S];
        Dft<T, A>::run(v, T(-1));
#pragma unroll
        for (int m = 0; m < A; ++m) x[m * S] = v[m];
    }
    __syncthreads();
    const double2* src = reinterpret_cast<const double2*>(psi + band * ldpsi * 2);
    double2* dst = reinterpret_cast<double2*>(hpsi + band * ldh * 2);
    const int j0 = a.stick_off[s0], j1 = a.stick_off[s0 + nsb];
    const int base = s0 * N;
    for (int j = j0 + tid; j < j1; j += nt) {
        const int p = a.ent_pw[j];
        const double t = a.ekin[p];
        const int pos = a.ent_pos[j] - base, sl = pos / N, z = pos - sl * N;
        const cx<T> v = buf[sl * lz + z + z / B];
        const double2 x = src[p];
        dst[p] = double2{t * x.x + static_cast<double>(v.x), t * x.y + static_cast<double>(v.y)};
    }
}

namespace {

// first-stage radix A and second-stage radix B of the square plane sizes with a kernel
struct TlShape {
    int n, a, b;
};
constexpr TlShape kTlShapes[] = {{24, 6, 4},  {30, 6, 5},   {32, 8, 4},   {36, 6, 6},   {40, 8, 5},
                                 {45, 9, 5},  {48, 8, 6},   {54, 9, 6},   {60, 10, 6},  {64, 8, 8},
                                 {72, 9, 8},  {75, 15, 5},  {80, 10, 8},  {81, 9, 9},   {90, 10, 9},
                                 {96, 12, 8}, {100, 10, 10}, {108, 12, 9}, {120, 12, 10}, {128, 16, 8},
                                 {135, 15, 9}, {144, 12, 12}};

int odd_at_least(int x) { return x | 1; }

}  // namespace

bool plane_tl_supported(int n1, int n2) {
    if (n1 != n2) return false;
    for (const TlShape& s : kTlShapes)
        if (s.n == n1) return true;
    return false;
}

PlaneTlConfig plane_tl_config(int n, int ncol, bool fp32, size_t max_lds, int num_cus) {
    PlaneTlConfig cfg{};
    int A = 0, B = 0;
    for (const TlShape& s : kTlShapes)
        if (s.n == n) A = s.a, B = s.b;
    if (A == 0) return cfg;
    const int G = plane_group_rows(A, B);  // rows per wave group (plane_tl)
    const size_t cs = fp32 ? 8 : 16;
    cfg.lc = odd_at_least(n + A);
    cfg.lr = cfg.lc;
    const size_t base = cs * plane_tw_entries(!fp32, A, B) + 4 * static_cast<size_t>((n + 3) & ~3);
    const size_t plane = cs * static_cast<size_t>(ncol) * cfg.lc;
    const size_t per_wave = cs * static_cast<size_t>(G) * cfg.lr;  // a wave's rows in flight
    // the plane goes to LDS when at least four waves' rows fit beside it
    cfg.plane_global = base + plane + 4 * per_wave > max_lds;
    const size_t fixed = base + (cfg.plane_global ? 0 : plane);
    if (fixed + per_wave > max_lds) return cfg;
    // waves: enough to cover the larger pass in one round of groups, up to 8 (512 threads keep
    // the FP64 radix-12/15/16 stages free of spills); two workgroups
    // per CU when the plane and four or more waves fit in half the LDS (the complex64 plane at
    // cfg2), which overlaps their passes; a plane kept in global memory instead runs 2x slower
    const int want = std::min(8, std::max(4, (std::max(n, ncol) + G - 1) / G));
    int nw = static_cast<int>(std::min<size_t>(want, (max_lds - fixed) / per_wave));
    const int nw_half = fixed < max_lds / 2 ? static_cast<int>((max_lds / 2 - fixed) / per_wave) : 0;
    const bool two_per_cu = !cfg.plane_global && nw_half >= 4;
    if (two_per_cu) nw = std::min(nw, nw_half);
    if (nw < 1) return cfg;
    cfg.rb = G;
    cfg.row_waves = nw;
    // one workgroup per CU: more waves (no row buffers of their own) until the column passes
    // take one round of groups, up to the kernel's launch bound
    const int max_waves = plane_tl_max_threads(!fp32, n) / 64;
    const int col_waves = two_per_cu ? nw : std::min(max_waves, std::max(nw, (ncol + G - 1) / G));
    cfg.threads = 64 * col_waves;
    cfg.lds = fixed + per_wave * nw;
    // FP64: persistent workgroups, one per CU; complex64: one per plane
    cfg.workgroups = fp32 ? std::numeric_limits<int>::max() : std::max(1, num_cus);
    cfg.plane_bytes = plane;
    cfg.ok = true;
    return cfg;
}

bool stick_tl_supported(int n3) {
    for (const TlShape& s : kTlShapes)
        if (s.n == n3) return true;
    return false;
}

StickTlConfig stick_tl_config(int n3, bool fp32) {
    StickTlConfig cfg{};
    int A = 0;
    for (const TlShape& s : kTlShapes)
        if (s.n == n3) A = s.a;
    if (A == 0) return cfg;
    const size_t cs = fp32 ? 8 : 16;
    cfg.lz = odd_at_least(n3 + A);
    // sticks per workgroup, from a sweep of 8..32 at cfg2 (96^3, 1,685 sticks): FP64 inverse 16
    // (0.149 against 0.167 ms at 32), forward 32; complex64 inverse 32, forward 8 (0.214 against
    // 0.244 ms)
    cfg.spb = fp32 ? 32 : 16;
    cfg.spb_f = fp32 ? 8 : 32;
    cfg.threads = 256;
    cfg.lds = cs * (n3 + static_cast<size_t>(cfg.spb) * cfg.lz);
    cfg.lds_f = cs * (n3 + static_cast<size_t>(cfg.spb_f) * cfg.lz);
    cfg.ok = true;
    return cfg;
}

template <typename T, int A, int B>
static hipError_t launch_sticks_shape(const HLocalArgs& a, int dir, const double* psi, long ldpsi, cx<T>* zbuf,
                                      double* hpsi, long ldh, const StickTlConfig& c, int nvec, hipStream_t st) {
    static std::atomic<bool> raised{false};
    if (!raised) {
        hipError_t e = hipFuncSetAttribute(reinterpret_cast<const void*>(&zsticks_inverse_tl<T, A, B>),
                                           hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024);
        if (e != hipSuccess) return e;
        e = hipFuncSetAttribute(reinterpret_cast<const void*>(&zsticks_forward_tl<T, A, B>),
                                hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024);
        if (e != hipSuccess) return e;
        raised = true;
    }
    const int spb = dir > 0 ? c.spb : c.spb_f;
    const int nblk = (a.nst + spb - 1) / spb;
    const dim3 grid(static_cast<unsigned>(8L * nblk * ((nvec + 7) / 8)));
    const cx<T>* tw = static_cast<const cx<T>*>(a.twp3);
    if (dir > 0)
        hipLaunchKernelGGL((zsticks_inverse_tl<T, A, B>), grid, dim3(c.threads), c.lds, st, a, psi, ldpsi, zbuf, tw,
                           spb, c.lz, nblk, nvec);
    else
        hipLaunchKernelGGL((zsticks_forward_tl<T, A, B>), grid, dim3(c.threads), c.lds_f, st, a, zbuf, psi, ldpsi,
                           hpsi, ldh, tw, spb, c.lz, nblk, nvec);
    return hipGetLastError();
}

template <typename T>
hipError_t launch_sticks_tl(const HLocalArgs& a, int dir, const double* psi, long ldpsi, cx<T>* zbuf, double* hpsi,
                            long ldh, const StickTlConfig& c, int nvec, hipStream_t st) {
    switch (a.p3.n) {
#define Q3X_TL(NN, AA, BB) \
    case NN: return launch_sticks_shape<T, AA, BB>(a, dir, psi, ldpsi, zbuf, hpsi, ldh, c, nvec, st);
        Q3X_TL(24, 6, 4)
        Q3X_TL(30, 6, 5)
        Q3X_TL(32, 8, 4)
        Q3X_TL(36, 6, 6)
        Q3X_TL(40, 8, 5)
        Q3X_TL(45, 9, 5)
        Q3X_TL(48, 8, 6)
        Q3X_TL(54, 9, 6)
        Q3X_TL(60, 10, 6)
        Q3X_TL(64, 8, 8)
        Q3X_TL(72, 9, 8)
        Q3X_TL(75, 15, 5)
        Q3X_TL(80, 10, 8)
        Q3X_TL(81, 9, 9)
        Q3X_TL(90, 10, 9)
        Q3X_TL(96, 12, 8)
        Q3X_TL(100, 10, 10)
        Q3X_TL(108, 12, 9)
        Q3X_TL(120, 12, 10)
        Q3X_TL(128, 16, 8)
        Q3X_TL(135, 15, 9)
        Q3X_TL(144, 12, 12)
#undef Q3X_TL
        default: return hipErrorInvalidValue;
    }
}

template hipError_t launch_sticks_tl<double>(const HLocalArgs&, int, const double*, long, cx<double>*, double*, long,
                                             const StickTlConfig&, int, hipStream_t);
template hipError_t launch_sticks_tl<float>(const HLocalArgs&, int, const double*, long, cx<float>*, double*, long,
                                            const StickTlConfig&, int, hipStream_t);

template <typename T, int A, int B, bool PG>
static hipError_t launch_one(const HLocalArgs& a, const T* vplane, cx<T>* zbuf, const PlaneTlConfig& c, int nvec,
                             void* pglob, hipStream_t st) {
    static std::atomic<bool> raised{false};
    if (!raised) {
        const hipError_t e = hipFuncSetAttribute(reinterpret_cast<const void*>(&plane_tl<T, A, B, PG>),
                                                 hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024);
        if (e != hipSuccess) return e;
        raised = true;
    }
    const int nplanes = a.p3.n * nvec;
    const int grid = std::min(nplanes, std::max(1, c.workgroups));
    hipLaunchKernelGGL((plane_tl<T, A, B, PG>), dim3(grid), dim3(c.threads), c.lds, st, a, vplane, zbuf,
                       static_cast<const cx<T>*>(a.twp), c.lc, c.lr, c.row_waves, nplanes, static_cast<cx<T>*>(pglob));
    return hipGetLastError();
}

template <typename T, int A, int B>
static hipError_t launch_shape(const HLocalArgs& a, const T* vplane, cx<T>* zbuf, const PlaneTlConfig& c, int nvec,
                               void* pglob, hipStream_t st) {
    if (c.plane_global) {
        if (!pglob) return hipErrorInvalidValue;
        return launch_one<T, A, B, true>(a, vplane, zbuf, c, nvec, pglob, st);
    }
    return launch_one<T, A, B, false>(a, vplane, zbuf, c, nvec, nullptr, st);
}

template <typename T>
hipError_t launch_plane_tl(const HLocalArgs& a, const T* vplane, cx<T>* zbuf, const PlaneTlConfig& c, int nvec,
                           void* pglob, hipStream_t st) {
    switch (a.p1.n) {
#define Q3X_TL(NN, AA, BB) \
    case NN: return launch_shape<T, AA, BB>(a, vplane, zbuf, c, nvec, pglob, st);
        Q3X_TL(24, 6, 4)
        Q3X_TL(30, 6, 5)
        Q3X_TL(32, 8, 4)
        Q3X_TL(36, 6, 6)
        Q3X_TL(40, 8, 5)
        Q3X_TL(45, 9, 5)
        Q3X_TL(48, 8, 6)
        Q3X_TL(54, 9, 6)
        Q3X_TL(60, 10, 6)
        Q3X_TL(64, 8, 8)
        Q3X_TL(72, 9, 8)
        Q3X_TL(75, 15, 5)
        Q3X_TL(80, 10, 8)
        Q3X_TL(81, 9, 9)
        Q3X_TL(90, 10, 9)
        Q3X_TL(96, 12, 8)
        Q3X_TL(100, 10, 10)
        Q3X_TL(108, 12, 9)
        Q3X_TL(120, 12, 10)
        Q3X_TL(128, 16, 8)
        Q3X_TL(135, 15, 9)
        Q3X_TL(144, 12, 12)
#undef Q3X_TL
        default: return hipErrorInvalidValue;
    }
}

template hipError_t launch_plane_tl<double>(const HLocalArgs&, const double*, cx<double>*, const PlaneTlConfig&, int,
                                            void*, hipStream_t);
template hipError_t launch_plane_tl<float>(const HLocalArgs&, const float*, cx<float>*, const PlaneTlConfig&, int,
                                           void*, hipStream_t);

// ---- charge density (PlanewaveChargeDensityAccumulator::accumulate, accumulator.hpp:77-99, and the
// per-band |psi(r)|^2 kernel of accumulator_gpu.hpp:267 / gpu_kernels_cuda.cu:156-172), FP64.
// Per (column, i3) plane of the stick buffer: the backward column and row transforms of
// plane_tl, then rho(i1, i2, i3) += w_re Re(f)^2 + w_im Im(f)^2 for the real-space values f of the
// column.  A column is one band (w_re = w_im = its weight) or, with the Gamma trick, a packed pair
// f = f1 + i f2 of real bands (w_re, w_im = their weights).  rho is the C-order (i1, i2, i3) box;
// the planes of different columns add into it atomically.
template <int A, int B>
__global__ void __launch_bounds__(512) plane_density(HLocalArgs a, const cx<double>* __restrict__ zbuf,
                                                     const cx<double>* __restrict__ twg, int lc, int lr, int nrw,
                                                     const double2* __restrict__ wts, double* __restrict__ rho) {
    using T = double;
    constexpr int N = A * B;
    constexpr int S = B + 1;
    constexpr int G = 64 / (A > B ? A : B);
    extern __shared__ __align__(16) unsigned char smem_raw[];
    cx<T>* tw = reinterpret_cast<cx<T>*>(smem_raw);
    int* colmap = reinterpret_cast<int*>(tw + N);
    cx<T>* P = reinterpret_cast<cx<T>*>(colmap + ((N + 3) & ~3));  // [ncol][lc]
    const int ncol = a.ncol, n3 = a.p3.n;
    const int tid = threadIdx.x, nt = blockDim.x;
    const int lane = tid & 63, wave = tid >> 6, nw = nt >> 6;
    cx<T>* Rw = P + ncol * lc + wave * G * lr;  // this wave's [G][lr] (waves < nrw)
    const int z = blockIdx.x, band = blockIdx.y;
    const cx<T> zero = {T(0), T(0)};
    for (int i = tid; i < N; i += nt) {
        tw[i] = twg[i];
        colmap[i] = -1;
    }
    for (int i = tid; i < ncol * lc; i += nt) P[i] = zero;
    __syncthreads();
    for (int c = tid; c < ncol; c += nt) colmap[a.col_i1[c]] = c;
    const cx<T>* zp = zbuf + (static_cast<long>(band) * n3 + z) * a.nst;
    for (int s = tid; s < a.nst; s += nt) {
        const int q = a.stick_i2[s];
        P[a.stick_col[s] * lc + q + q / B] = zp[s];
    }
    __syncthreads();
    const int gB = lane / B, iB = lane - gB * B;
    const int gA = lane / A, iA = lane - gA * A;
    // columns, backward (as plane_tl)
    for (int c0 = wave * G; c0 < ncol; c0 += nw * G) {
        if (gB < G && c0 + gB < ncol) {
            cx<T>* col = P + (c0 + gB) * lc + iB;
            cx<T> v[A];
#pragma unroll
            for (int j = 0; j < A; ++j) v[j] = col[j * S];
            Dft<T, A>::run(v, T(1));
#pragma unroll
            for (int k = 1; k < A; ++k) v[k] = cmulc(v[k], tw[iB * k]);
#pragma unroll
            for (int k = 0; k < A; ++k) col[k * S] = v[k];
        }
        wave_order();
        if (gA < G && c0 + gA < ncol) {
            cx<T>* blk = P + (c0 + gA) * lc + iA * S;
            cx<T> u[B];
#pragma unroll
            for (int j = 0; j < B; ++j) u[j] = blk[j];
            Dft<T, B>::run(u, T(1));
#pragma unroll
            for (int j = 0; j < B; ++j) blk[j] = u[j];
        }
        wave_order();
    }
    __syncthreads();
    const double wr = wts[band].x, wi = wts[band].y;
    // rows, backward, then the weighted squares; row i2 sits at slot (i2 % A) * S + i2 / A
    for (int r0 = wave * G; wave < nrw && r0 < N; r0 += nrw * G) {
        if (gB < G && r0 + gB < N) {
            const int i2 = r0 + gB;
            const int pp = (i2 % A) * S + i2 / A;
            cx<T> v[A];
#pragma unroll
            for (int j = 0; j < A; ++j) {
                const int c = colmap[B * j + iB];
                v[j] = zero;
                if (c >= 0) v[j] = P[c * lc + pp];
            }
            Dft<T, A>::run(v, T(1));
#pragma unroll
            for (int k = 1; k < A; ++k) v[k] = cmulc(v[k], tw[iB * k]);
            cx<T>* row = Rw + gB * lr + iB;
#pragma unroll
            for (int k = 0; k < A; ++k) row[k * S] = v[k];
        }
        wave_order();
        if (gA < G && r0 + gA < N) {  // lane k1: elements i1 = k1 + A j of row i2
            const int i2 = r0 + gA;
            const cx<T>* blk = Rw + gA * lr + iA * S;
            cx<T> u[B];
#pragma unroll
            for (int j = 0; j < B; ++j) u[j] = blk[j];
            Dft<T, B>::run(u, T(1));
#pragma unroll
            for (int j = 0; j < B; ++j) {
                const long i1 = iA + A * j;
                atomicAdd(rho + (i1 * N + i2) * n3 + z, wr * u[j].x * u[j].x + wi * u[j].y * u[j].y);
            }
        }
        wave_order();
    }
}

template <int A, int B>
static hipError_t launch_density_shape(const HLocalArgs& a, const cx<double>* zbuf, const PlaneTlConfig& c, int nvec,
                                       const double* wts, double* rho, hipStream_t st) {
    static std::atomic<bool> raised{false};
    if (!raised) {
        const hipError_t e = hipFuncSetAttribute(reinterpret_cast<const void*>(&plane_density<A, B>),
                                                 hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024);
        if (e != hipSuccess) return e;
        raised = true;
    }
    hipLaunchKernelGGL((plane_density<A, B>), dim3(a.p3.n, nvec), dim3(std::min(c.threads, 512)), c.lds, st, a, zbuf,
                       static_cast<const cx<double>*>(a.twp), c.lc, c.lr, c.row_waves,
                       reinterpret_cast<const double2*>(wts), rho);
    return hipGetLastError();
}

hipError_t launch_plane_density(const HLocalArgs& a, const cx<double>* zbuf, const PlaneTlConfig& c, int nvec,
                                const double* wts, double* rho, hipStream_t st) {
    if (!c.ok || c.plane_global) return hipErrorInvalidValue;
    switch (a.p1.n) {
#define Q3X_TL(NN, AA, BB) \
    case NN: return launch_density_shape<AA, BB>(a, zbuf, c, nvec, wts, rho, st);
        Q3X_TL(24, 6, 4)
        Q3X_TL(30, 6, 5)
        Q3X_TL(32, 8, 4)
        Q3X_TL(36, 6, 6)
        Q3X_TL(40, 8, 5)
        Q3X_TL(45, 9, 5)
        Q3X_TL(48, 8, 6)
        Q3X_TL(54, 9, 6)
        Q3X_TL(60, 10, 6)
        Q3X_TL(64, 8, 8)
        Q3X_TL(72, 9, 8)
        Q3X_TL(75, 15, 5)
        Q3X_TL(80, 10, 8)
        Q3X_TL(81, 9, 9)
        Q3X_TL(90, 10, 9)
        Q3X_TL(96, 12, 8)
        Q3X_TL(100, 10, 10)
        Q3X_TL(108, 12, 9)
        Q3X_TL(120, 12, 10)
        Q3X_TL(128, 16, 8)
        Q3X_TL(135, 15, 9)
        Q3X_TL(144, 12, 12)
#undef Q3X_TL
        default: return hipErrorInvalidValue;
    }
}

}  // namespace q3x
