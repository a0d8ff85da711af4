// Generalized Hermitian eigensolver H x = w S x for the Davidson subspace (M <= a few hundred),
// entirely on the device.  zhegv itype 1 semantics (SerialSubspaceDiagonalizer::solve,
// subspace/serial.hpp:26-136; the GPU interface math/linear_algebra/gpu/lapack/init.hpp:77-82),
// computed by a different route that suits one MI355X:
//
//   1. Z = S^(-1/2) (Loewdin).  The Davidson basis is orthonormalised, so S = I + E with tiny E
//      and Z = 1.5 I - 0.5 S is exact to O(|E|^2).  Otherwise a coupled Newton-Schulz
//      iteration on the MFMA GEMM converges to S^(-1/2); it fails to converge exactly when S is
//      not positive definite, which is reported as zhegv's info > n.
//   2. C = Z H Z (two GEMMs).  The eigenvectors of the pencil are X = Z Y with C Y = Y w, and
//      X^H S X = I.
//   3. C = Q T Q^H by Householder reflectors in one workgroup: the rank-2 update of step k and
//      the matrix-vector product of step k+1 share one pass over the trailing block (look-
//      ahead on the next column), so each step reads and writes the block once.  For n <= 99
//      the matrix sits in LDS.
//   4. T = D Tr D^H with a real tridiagonal Tr (phases D); eigenvalues of Tr by multisection:
//      one wavefront per eigenvalue, 63 Sturm counts per round, 6 bits per round.
//   5. Eigenvectors of Tr, one thread per vector, in parallel: an isolated eigenvalue by one
//      twisted factorisation (the core step of MRRR); eigenvalues closer than 1e-6 |T| by
//      inverse iteration, orthogonalised against each other in rounds.
//   6. Y = Q D Z_tr (reflectors applied per column, one wavefront each), Y <- Y (1.5 I - 0.5
//      Y^H Y) to restore orthonormality to rounding, X = Z Y.
#include "subspace_eig.h"

#include <hip/hip_runtime.h>

#include <atomic>
#include <cstring>

#include <algorithm>
#include <cfloat>
#include <cmath>
#include <vector>

#include "zgemm.h"

namespace q3x {
namespace {

struct z2 {
    double x, y;
};
__device__ __forceinline__ z2 zmul(z2 a, z2 b) { return {a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x}; }
// a * conj(b)
__device__ __forceinline__ z2 zmulc(z2 a, z2 b) { return {a.x * b.x + a.y * b.y, a.y * b.x - a.x * b.y}; }

constexpr int kThreads = 1024;  // single-workgroup kernels: 16 wavefronts
constexpr double kEps = 2.220446049250313e-16;

// Cross-lane move of a double by a DPP pattern (each 32-bit half separately).
template <int CTRL>
__device__ __forceinline__ double dpp_mov(double v) {
    const int lo = __builtin_amdgcn_mov_dpp(__double2loint(v), CTRL, 0xF, 0xF, false);
    const int hi = __builtin_amdgcn_mov_dpp(__double2hiint(v), CTRL, 0xF, 0xF, false);
    return __hiloint2double(hi, lo);
}
// a / b to about an ulp: the hardware reciprocal refined by two Newton steps, then one product.
// The recurrences that use it (Householder reflectors, Sturm counts, twisted factorisations)
// need each operation's relative error to be O(eps), not a correctly rounded quotient, and
// this chain is shorter than the full division sequence on their serial paths.
__device__ __forceinline__ double div_ulp(double a, double b) {
    double r = __builtin_amdgcn_rcp(b);
    r = fma(r, fma(-b, r, 1.0), r);
    r = fma(r, fma(-b, r, 1.0), r);
    return a * r;
}

__device__ __forceinline__ double read_lane(double v, int lane) {
    return __hiloint2double(__builtin_amdgcn_readlane(__double2hiint(v), lane),
                            __builtin_amdgcn_readlane(__double2loint(v), lane));
}
// Sum over the 64 lanes (all active), the same value in every lane: within each row of 16
// lanes by DPP rotations (row_ror 8, row_ror 4, quad_perm [2,3,0,1], quad_perm [1,0,3,2]),
// then the four row sums read from lanes 0, 16, 32, 48 in a fixed order.  DPP moves stay in the
// VALU; a shuffle through ds_bpermute costs an LDS round trip per step.
__device__ __forceinline__ double wave_sum(double v) {
    v += dpp_mov<0x128>(v);
    v += dpp_mov<0x124>(v);
    v += dpp_mov<0x4E>(v);
    v += dpp_mov<0xB1>(v);
    return (read_lane(v, 0) + read_lane(v, 16)) + (read_lane(v, 32) + read_lane(v, 48));
}
// Sums of V values per lane over the wavefront at once (V a power of 2, <= 64): each exchange
// stage swaps half of the live values with the partner lane and adds, halving V, so V sums
// cost about log2(64) + V - 1 exchanges instead of 6 V.  On return lane l holds the full sum of
// value l / (64 / V).
// Orders a wavefront's LDS accesses across lanes (they execute in issue order; this keeps the
// compiler from moving them across the point).
__device__ __forceinline__ void wave_fence() {
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    __builtin_amdgcn_wave_barrier();
}

template <int V>
__device__ __forceinline__ double wave_multi_sum(double (&x)[V], int lane) {
    int o = 32;
#pragma unroll
    for (int h = V / 2; h >= 1; h /= 2, o /= 2) {
        const bool up = (lane & o) != 0;
#pragma unroll
        for (int i = 0; i < h; ++i) {
            // both values in registers before the select: a select between two loads of x is
            // otherwise folded into one load at a lane-dependent index, which moves x to scratch
            double lo = x[i], hi = x[i + h];
            asm volatile("" : "+v"(lo), "+v"(hi));
            const double send = up ? lo : hi;
            const double keep = up ? hi : lo;
            x[i] = keep + __shfl_xor(send, o, 64);
        }
    }
    // the remaining plain sums: across rows by shuffles, within a row of 16 by DPP moves (row
    // mirror, half-row mirror, quad permutes: each pairs the lanes of the group still to be
    // summed, in the VALU, without an LDS round trip)
    double s = x[0];
#pragma unroll
    for (; o >= 16; o /= 2) s += __shfl_xor(s, o, 64);
    if (o >= 8) s += dpp_mov<0x140>(s);
    if (o >= 4) s += dpp_mov<0x141>(s);
    if (o >= 2) s += dpp_mov<0x4E>(s);
    if (o >= 1) s += dpp_mov<0xB1>(s);
    return s;
}
__device__ __forceinline__ double wave_max(double v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v = fmax(v, __shfl_xor(v, o, 64));
    return v;
}
__device__ __forceinline__ double wave_min(double v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v = fmin(v, __shfl_xor(v, o, 64));
    return v;
}

// Block reductions.  red holds 2 x 16 slots used alternately, so one barrier per reduction
// suffices: a slot is rewritten only after every thread has passed the next reduction's
// barrier, i.e. after it finished reading the slot.  Every thread gets the same result
// (same summation order).
__device__ double block_sum(double v, double* red, int& slot) {
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6, nw = blockDim.x >> 6;
    v = wave_sum(v);
    double* r = red + slot * 16;
    if (lane == 0) r[w] = v;
    __syncthreads();
    double s = 0.0;
    for (int i = 0; i < nw; ++i) s += r[i];
    slot ^= 1;
    return s;
}
__device__ double block_max(double v, double* red, int& slot) {
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6, nw = blockDim.x >> 6;
    v = wave_max(v);
    double* r = red + slot * 16;
    if (lane == 0) r[w] = v;
    __syncthreads();
    double s = r[0];
    for (int i = 1; i < nw; ++i) s = fmax(s, r[i]);
    slot ^= 1;
    return s;
}

// ---------------------------------------------------------------- 1. preparation ----
// Hf, Sf: full Hermitian n x n (ld n) from the upper triangles; Zl = 1.5 I - 0.5 Sf;
// *err_bits = bits of max |S - I| (componentwise; non-negative doubles order like their bit
// patterns, so the maximum over workgroups is an integer atomic max; *err_bits starts at 0).
__global__ __launch_bounds__(256) void herm_prep(int n, const z2* __restrict__ H, long ldh,
                                                 const z2* __restrict__ S, long lds, z2* __restrict__ Hf,
                                                 z2* __restrict__ Sf, z2* __restrict__ Zl,
                                                 unsigned long long* err_bits) {
    __shared__ double red[4];
    double dev = 0.0;
    const long nn = static_cast<long>(n) * n;
    for (long e = blockIdx.x * 256L + threadIdx.x; e < nn; e += 256L * gridDim.x) {
        const int i = static_cast<int>(e % n), j = static_cast<int>(e / n);
        z2 h, s;
        if (i <= j) {
            h = H[i + j * ldh];
            s = S[i + j * lds];
        } else {
            const z2 ht = H[j + i * ldh], st = S[j + i * lds];
            h = {ht.x, -ht.y};
            s = {st.x, -st.y};
        }
        if (i == j) {
            h.y = 0.0;
            s.y = 0.0;
            dev = fmax(dev, fabs(s.x - 1.0));
            Zl[e] = {1.5 - 0.5 * s.x, 0.0};
        } else {
            dev = fmax(dev, fmax(fabs(s.x), fabs(s.y)));
            Zl[e] = {-0.5 * s.x, -0.5 * s.y};
        }
        Hf[e] = h;
        Sf[e] = s;
    }
    if (dev != dev) dev = INFINITY;
    dev = wave_max(dev);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = dev;
    __syncthreads();
    if (threadIdx.x == 0) {
        const double m = fmax(fmax(red[0], red[1]), fmax(red[2], red[3]));
        atomicMax(err_bits, static_cast<unsigned long long>(__double_as_longlong(m)));
    }
}

// Y = Sf / c, Z = I
__global__ void ns_init(int n, const z2* __restrict__ Sf, double inv_c, z2* __restrict__ Y, z2* __restrict__ Z) {
    const long nn = static_cast<long>(n) * n;
    for (long e = blockIdx.x * static_cast<long>(blockDim.x) + threadIdx.x; e < nn;
         e += static_cast<long>(gridDim.x) * blockDim.x) {
        const z2 s = Sf[e];
        Y[e] = {s.x * inv_c, s.y * inv_c};
        Z[e] = {(e % n == e / n) ? 1.0 : 0.0, 0.0};
    }
}

// T = 1.5 I - 0.5 ZY ; err[0] = max |I - ZY|.  One workgroup.
__global__ __launch_bounds__(kThreads) void ns_residual(int n, const z2* __restrict__ ZY, z2* __restrict__ T,
                                                        double* err) {
    __shared__ double red[32];
    int slot = 0;
    double dev = 0.0;
    const long nn = static_cast<long>(n) * n;
    for (long e = threadIdx.x; e < nn; e += blockDim.x) {
        const z2 a = ZY[e];
        const bool diag = (e % n) == (e / n);
        const double r = diag ? 1.0 - a.x : -a.x, im = -a.y;
        dev = fmax(dev, fmax(fabs(r), fabs(im)));
        T[e] = {(diag ? 1.5 : 0.0) - 0.5 * a.x, -0.5 * a.y};
    }
    // NaN-propagating: a NaN anywhere makes the result NaN
    const double m = block_max(dev != dev ? INFINITY : dev, red, slot);
    if (threadIdx.x == 0) err[0] = m;
}

// A *= s (n x n)
__global__ void scale_matrix(long nn, double s, z2* __restrict__ A) {
    for (long e = blockIdx.x * static_cast<long>(blockDim.x) + threadIdx.x; e < nn;
         e += static_cast<long>(gridDim.x) * blockDim.x) {
        const z2 a = A[e];
        A[e] = {a.x * s, a.y * s};
    }
}

// ---------------------------------------------------------------- 3. tridiagonalisation ----
// Reflector k: H_k = I - tau v v^H on rows [k+1, n), v_0 = 1, H_k^H x = beta e_0 for
// x = column k rows [k+1, n).  Its vector is kept in column k rows k+2.. below the implicit 1.
//
// Step k applies the rank-2 update B -= v q^H + q v^H (q = p - (tau/2)(v^H p) v, p = tau B v)
// to the trailing block.  The update of step k and the product p of step k+1 share one pass
// over the block; column k+2 (whose reflector step k+2 needs) is finished by wavefront 0 while
// the others form q.  Per step: one pass and three barriers.

// Column access of the tridiagonalisation.  For the L2-resident matrix, loads and stores go
// through a buffer resource bounded to the column's live rows: rows past the end read as 0 and
// their stores are dropped by the hardware, so the row loops need no per-lane branches (with
// branches the compiler serialises every store behind a full vmcnt wait).  The LDS-resident
// matrix uses plain accesses with the same guard.
typedef unsigned u32x4 __attribute__((ext_vector_type(4)));
template <bool LM>
struct ColumnView {
    z2* base;
    int len;
    __amdgpu_buffer_rsrc_t rs;
    __device__ ColumnView(z2* b, int l) : base(b), len(l) {
        if (!LM) rs = __builtin_amdgcn_make_buffer_rsrc(b, 0, l * 16, 0x00020000);
    }
    __device__ z2 load(int i) const {
        if (LM) return i < len ? base[i] : z2{0.0, 0.0};
        const u32x4 v = __builtin_amdgcn_raw_buffer_load_b128(rs, i * 16, 0, 0);
        z2 r;
        __builtin_memcpy(&r, &v, 16);
        return r;
    }
    __device__ void store(int i, z2 a) const {
        if (LM) {
            if (i < len) base[i] = a;
            return;
        }
        u32x4 v;
        __builtin_memcpy(&v, &a, 16);
        __builtin_amdgcn_raw_buffer_store_b128(v, rs, i * 16, 0, 0);
    }
};

// Wavefront 0: finish column j (rows [j, n)) with reflector j-1 (vn, q = p - f vn, if tn != 0),
// record d[j], then form reflector j into vout (and column j's storage) when j <= n-3.
// Returns tau_j (0 when there is none).  Lane l holds rows j + l + 64 t.
template <int RPL, bool LM>
__device__ double column_step(z2* M, int n, int j, double tn, const z2* vn, const z2* p, double f, z2* vout,
                              double* d, z2* beta, double* tau) {
    const int lane = threadIdx.x & 63;
    const int L = n - j;
    const ColumnView<LM> col(M + j + static_cast<long>(j) * n, L);
    z2 a[RPL];
#pragma unroll
    for (int t = 0; t < RPL; ++t) a[t] = col.load(lane + 64 * t);
    if (tn != 0.0) {
        const z2 v0 = vn[0];
        const z2 q0 = {p[0].x - f * v0.x, p[0].y - f * v0.y};
#pragma unroll
        for (int t = 0; t < RPL; ++t) {
            const int i = lane + 64 * t;
            const int ic = min(i, L - 1);  // LDS vectors hold L entries
            const z2 vi = vn[ic];
            const z2 qi = {p[ic].x - f * vi.x, p[ic].y - f * vi.y};
            const z2 t1 = zmulc(vi, q0), t2 = zmulc(qi, v0);
            a[t].x -= t1.x + t2.x;
            a[t].y -= t1.y + t2.y;
            col.store(i, a[t]);
            if (i >= L) a[t] = {0.0, 0.0};
        }
    }
    if (lane == 0) d[j] = a[0].x;
    const z2 x0 = {__shfl(a[0].x, 1, 64), __shfl(a[0].y, 1, 64)};  // row j+1
    if (j == n - 2) {
        if (lane == 0) beta[n - 1] = x0;
        return 0.0;
    }
    if (j > n - 3) return 0.0;
    double part = 0.0;
#pragma unroll
    for (int t = 0; t < RPL; ++t) {
        const int i = lane + 64 * t;
        if (i >= 2) part += a[t].x * a[t].x + a[t].y * a[t].y;  // rows past the end hold 0
    }
    const double tail = wave_sum(part);
    double tj = 0.0;
    if (tail == 0.0) {
        if (lane == 0) beta[j + 1] = x0;
    } else {
        const double a0 = hypot(x0.x, x0.y);
        const double xnorm = sqrt(tail + a0 * a0);
        const double ia0 = a0 == 0.0 ? 0.0 : div_ulp(1.0, a0);
        const z2 ph = a0 == 0.0 ? z2{1.0, 0.0} : z2{x0.x * ia0, x0.y * ia0};
        const z2 b = {-ph.x * xnorm, -ph.y * xnorm};
        const z2 u0 = {x0.x - b.x, x0.y - b.y};
        const double iu2 = div_ulp(1.0, u0.x * u0.x + u0.y * u0.y);
        tj = div_ulp(2.0, 1.0 + tail * iu2);
#pragma unroll
        for (int t = 0; t < RPL; ++t) {
            const int i = lane + 64 * t;
            const z2 w = i == 1 ? z2{1.0, 0.0}
                                : z2{(a[t].x * u0.x + a[t].y * u0.y) * iu2, (a[t].y * u0.x - a[t].x * u0.y) * iu2};
            if (i >= 1 && i < L) vout[i - 1] = w;
            if (i >= 2) col.store(i, w);
        }
        if (lane == 0) beta[j + 1] = b;
    }
    if (lane == 0) tau[j] = tj;
    return tj;
}

// Pass of step k over columns / rows [k+2, n): rank-2 update with (v, q) when tk != 0, then
// p[c] = tn * (column)^H vn when tn != 0.  Wavefront w takes CB adjacent columns at a time.  The
// row vectors v, q, vn are the same for every column of the pass, so each lane loads its rows of
// them into registers once (they share LDS with the matrix, so re-reading them after a column
// store would be a dependent LDS round trip per column); the CB columns' loads are issued
// together, and their 2 CB wave reductions run side by side with no branch between them.
// v, q index rows from k+1; vn and p from k+2.
template <int RPL, int CB, bool LM>
__device__ void fused_pass(z2* M, int n, int k, double tk, const z2* v, const z2* q, double tn, const z2* vn,
                           z2* p) {
    const int lane = threadIdx.x & 63;
    const int wv = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6), nw = blockDim.x >> 6;
    const int len = n - k - 2, base = k + 2;
    if (wv * CB >= len) return;
    z2 vr[RPL], qr[RPL], wr[RPL];
    double msk[RPL];
#pragma unroll
    for (int t = 0; t < RPL; ++t) {
        const int r = lane + 64 * t;
        const int rc = min(r, len - 1);  // LDS vectors: stay in range; rows past the end are 0
        vr[t] = tk != 0.0 ? v[rc + 1] : z2{0.0, 0.0};
        qr[t] = tk != 0.0 ? q[rc + 1] : z2{0.0, 0.0};
        wr[t] = vn[rc];
        msk[t] = r < len ? 1.0 : 0.0;
    }
    for (int c0 = wv * CB; c0 < len; c0 += nw * CB) {
        z2 a[CB][RPL], qj[CB], vj[CB];
#pragma unroll
        for (int b = 0; b < CB; ++b) {
            const int c = min(c0 + b, len - 1);  // a missing column re-reads the last one (unused)
            const ColumnView<LM> col(M + static_cast<long>(base + c) * n + base, len);
#pragma unroll
            for (int t = 0; t < RPL; ++t) a[b][t] = col.load(lane + 64 * t);
            qj[b] = tk != 0.0 ? q[c + 1] : z2{0.0, 0.0};
            vj[b] = tk != 0.0 ? v[c + 1] : z2{0.0, 0.0};
        }
        double sr[CB], si[CB];
#pragma unroll
        for (int b = 0; b < CB; ++b) {
            sr[b] = 0.0;
            si[b] = 0.0;
#pragma unroll
            for (int t = 0; t < RPL; ++t) {
                if (tk != 0.0) {
                    const z2 t1 = zmulc(vr[t], qj[b]), t2 = zmulc(qr[t], vj[b]);
                    a[b][t].x -= t1.x + t2.x;
                    a[b][t].y -= t1.y + t2.y;
                }
                sr[b] += msk[t] * (a[b][t].x * wr[t].x + a[b][t].y * wr[t].y);
                si[b] += msk[t] * (a[b][t].x * wr[t].y - a[b][t].y * wr[t].x);
            }
        }
        if (tk != 0.0) {
#pragma unroll
            for (int b = 0; b < CB; ++b) {
                if (c0 + b < len) {
                    const ColumnView<LM> col(M + static_cast<long>(base + c0 + b) * n + base, len);
#pragma unroll
                    for (int t = 0; t < RPL; ++t) col.store(lane + 64 * t, a[b][t]);
                }
            }
        }
        if (tn != 0.0) {
            double x[2 * CB];
#pragma unroll
            for (int b = 0; b < CB; ++b) {
                x[2 * b] = sr[b];
                x[2 * b + 1] = si[b];
            }
            const double tot = wave_multi_sum<2 * CB>(x, lane);
            constexpr int span = 64 / (2 * CB);
            const int idx = lane / span, b = idx >> 1;
            if (lane % span == 0 && c0 + b < len) reinterpret_cast<double*>(p + c0 + b)[idx & 1] = tn * tot;
        }
    }
}

// A (n x n, ld n, full storage, Hermitian up to rounding) -> d (diagonal), beta (sub-diagonal,
// beta[i] couples i-1 and i), tau (reflectors, kept in A below the sub-diagonal), the real
// tridiagonal |beta| in eabs and the phases D in phase (T = D Tr D^H).  RPL >= ceil(n / 64).
template <bool LM, int RPL>
__global__ __launch_bounds__(kThreads) void tridiag_kernel(z2* A, int n, double* d, z2* beta, double* tau,
                                                           z2* phase, double* eabs) {
    constexpr int CB = RPL <= 2 ? 4 : (RPL <= 4 ? 2 : 1);
    extern __shared__ z2 sh[];
    z2* vb0 = sh;
    z2* vb1 = sh + n;
    z2* q = sh + 2 * n;
    z2* p = sh + 3 * n;
    double& tnext = *reinterpret_cast<double*>(sh + 4 * n + 16);  // sh + 4n .. 4n + 15: spare
    z2* M = LM ? sh + 4 * n + 17 : A;
    // d, beta and tau are kept in LDS and written to global memory once at the end
    z2* sbeta = sh + 4 * n + 17 + (LM ? static_cast<long>(n) * n : 0);
    double* sd = reinterpret_cast<double*>(sbeta + n);
    double* stau = sd + n;
    const int tid = threadIdx.x, nt = blockDim.x, wv = tid >> 6;
    const long nn = static_cast<long>(n) * n;
    if (LM) {
        for (long e = tid; e < nn; e += nt) M[e] = A[e];
        __syncthreads();
    }
    // Hermitian part: (A + A^H) / 2, real diagonal
    for (long e = tid; e < nn; e += nt) {
        const int i = static_cast<int>(e % n), j = static_cast<int>(e / n);
        if (i < j) {
            const z2 a = M[e], b = M[j + static_cast<long>(i) * n];
            const z2 h = {0.5 * (a.x + b.x), 0.5 * (a.y - b.y)};
            M[e] = h;
            M[j + static_cast<long>(i) * n] = {h.x, -h.y};
        } else if (i == j) {
            M[e].y = 0.0;
        }
    }
    for (int i = tid; i < n; i += nt) {
        sbeta[i] = {0.0, 0.0};
        stau[i] = 0.0;
        sd[i] = 0.0;
    }
    __syncthreads();
    if (n <= 2) {
        if (tid == 0) {
            sd[0] = M[0].x;
            if (n == 2) {
                sd[1] = M[3].x;
                sbeta[1] = M[1];
            }
        }
    } else {
        // reflector 0 (column 0 needs no update)
        if (wv == 0) {
            const double t0 = column_step<RPL, LM>(M, n, 0, 0.0, nullptr, nullptr, 0.0, vb0, sd, sbeta, stau);
            if ((tid & 63) == 0) tnext = t0;
        }
        __syncthreads();
        double tk = 0.0, tn = tnext;
        z2* v = vb1;   // reflector k (not yet defined at k = -1)
        z2* vn = vb0;  // reflector k+1
        // Once the trailing block has at most 64 rows, one row per lane covers it: the passes and
        // column steps drop the second row set, which would only add masked zeros.
        for (int k = -1; k <= n - 3; ++k) {
            const bool one_row = RPL > 1 && n - k - 2 <= 64;
            if (tk != 0.0 || tn != 0.0) {
                if (one_row)
                    fused_pass<1, 4, LM>(M, n, k, tk, v, q, tn, vn, p);
                else
                    fused_pass<RPL, CB, LM>(M, n, k, tk, v, q, tn, vn, p);
            }
            __syncthreads();
            const int len = n - k - 2;
            if (wv == 0) {  // f = tn vn^H p / 2, q = p - f vn, next column and reflector: one wavefront
                const int lane = tid & 63;
                double f = 0.0;
                if (tn != 0.0) {
                    double part = 0.0;
                    for (int i = lane; i < len; i += 64) part += vn[i].x * p[i].x + vn[i].y * p[i].y;
                    f = 0.5 * tn * wave_sum(part);
                    for (int i = lane; i < len; i += 64) q[i] = {p[i].x - f * vn[i].x, p[i].y - f * vn[i].y};
                }
                const double tj = one_row ? column_step<1, LM>(M, n, k + 2, tn, vn, p, f, v, sd, sbeta, stau)
                                          : column_step<RPL, LM>(M, n, k + 2, tn, vn, p, f, v, sd, sbeta, stau);
                if (lane == 0) tnext = tj;
            }
            __syncthreads();
            tk = tn;
            tn = tnext;
            z2* t = v;
            v = vn;
            vn = t;
        }
    }
    __syncthreads();
    for (int i = tid; i < n; i += nt) {
        d[i] = sd[i];
        beta[i] = sbeta[i];
        tau[i] = stau[i];
    }
    if (tid == 0) {
        z2 ph = {1.0, 0.0};
        phase[0] = ph;
        eabs[0] = 0.0;
        for (int i = 1; i < n; ++i) {
            const z2 b = sbeta[i];
            const double r = hypot(b.x, b.y);
            eabs[i] = r;
            if (r != 0.0) ph = zmul(ph, z2{b.x / r, b.y / r});
            phase[i] = ph;
        }
    }
    if (LM)
        for (long e = tid; e < nn; e += nt) A[e] = M[e];
}

// ---------------------------------------------------------------- 3b. tridiagonalisation over several CUs ----
// The same reduction with the matrix spread over the LDS of W workgroups: column j lives in
// workgroup j mod W (all rows), so the passes run at LDS speed on W CUs instead of at one CU's
// L2 bandwidth.  Per step: every workgroup updates its own columns and forms its entries of p;
// the owner of the next pivot column also publishes that column.  One grid barrier; then every
// workgroup reads p and the pivot column and forms q and the next reflector itself -- the same
// operations in the same order in each workgroup, so all hold identical values and nothing else
// has to be broadcast.  p and the pivot column are double-buffered by step parity: a workgroup
// can be at most one step ahead of another.  Workgroup 0 writes the outputs (d, beta, tau and the
// reflectors into R, n x n, column k rows k+2..).  At most one workgroup per CU, so all are
// resident on an otherwise idle device; the barrier's wait is bounded and sets *fail instead of
// waiting forever.

// Once any workgroup has given up (about a second of polling), every later barrier returns at
// once, so the kernel always drains quickly; the host turns *fail into an error.
// Each workgroup owns one flag word (flags[32 w], a 128-byte line of its own) and stores the
// step number there; its first wavefront polls the other W - 1 flags, one per lane.  Stores to
// separate lines replace W atomic adds to one word, which the memory side serialises: at
// n = 320 (20 workgroups) a reduction takes 4.8% less time, at n = 640 (64) 15% less.
// Release fence, relaxed store, relaxed polls, one acquire fence: each fence pair synchronises
// through the flag it reads, and no poll pays for an acquire (an agent-scope acquire
// invalidates cache lines; a seq_cst fence on each side would write back and invalidate twice).
__device__ void grid_barrier(unsigned* flags, int W, unsigned step, int* fail) {
    __syncthreads();
    if (threadIdx.x < 64) {
        const int lane = threadIdx.x, me = blockIdx.x;
        if (lane == 0) {
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");  // this workgroup's global writes
            __hip_atomic_store(flags + 32 * me, step, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        }
        unsigned spins = 0;
        for (;;) {
            bool here = true;
            for (int w = lane; w < W; w += 64)
                if (w != me) here = here && __hip_atomic_load(flags + 32 * w, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) >= step;
            if (__all(here)) break;
            const int f = __builtin_amdgcn_readfirstlane(__hip_atomic_load(fail, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT));
            if (f != 0 || ++spins > (1u << 20)) {
                if (lane == 0) atomicExch(fail, 1);
                break;
            }
            __builtin_amdgcn_s_sleep(1);
        }
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");  // the other workgroups' writes
    }
    __syncthreads();
}

// Block-level reflector of x = col[1..L) (col: LDS, length L = n - j, rows j..n-1 of column j).
// Writes v (length L-1) to vout, returns tau; workgroup 0 stores beta[j+1], tau[j] and v[1..]
// into R.  Barriers inside; every thread returns the same tau.
__device__ double block_house(const z2* col, int L, int n, int j, z2* vout, double* red, int& slot, bool writer,
                              z2* beta, double* tau, z2* R) {
    const int tid = threadIdx.x, nt = blockDim.x;
    double part = 0.0;
    for (int i = tid + 2; i < L; i += nt) part += col[i].x * col[i].x + col[i].y * col[i].y;
    const double tail = block_sum(part, red, slot);
    const z2 x0 = col[1];
    double t = 0.0;
    if (tail == 0.0) {
        if (writer && tid == 0) beta[j + 1] = x0;
    } else {
        const double a0 = hypot(x0.x, x0.y);
        const double xnorm = sqrt(tail + a0 * a0);
        const z2 ph = a0 == 0.0 ? z2{1.0, 0.0} : z2{x0.x / a0, x0.y / a0};
        const z2 b = {-ph.x * xnorm, -ph.y * xnorm};
        const z2 u0 = {x0.x - b.x, x0.y - b.y};
        const double iu2 = 1.0 / (u0.x * u0.x + u0.y * u0.y);
        t = 2.0 / (1.0 + tail * iu2);
        for (int i = tid + 1; i < L; i += nt) {
            const z2 a = col[i];
            const z2 w = i == 1 ? z2{1.0, 0.0} : z2{(a.x * u0.x + a.y * u0.y) * iu2, (a.y * u0.x - a.x * u0.y) * iu2};
            vout[i - 1] = w;
            if (writer && i >= 2) R[(j + i) + static_cast<long>(j) * n] = w;
        }
        if (writer && tid == 0) beta[j + 1] = b;
    }
    if (writer && tid == 0) tau[j] = t;
    __syncthreads();
    return t;
}

// Wavefront-level block_house: the calling wavefront alone forms the reflector of
// x = col[1..L) into vout (length L-1) and returns tau; workgroup 0's wavefront stores beta[j+1],
// tau[j] and v[1..] into R.  No block barrier.
__device__ double wave_house(const z2* col, int L, int n, int j, z2* vout, bool writer, z2* beta, double* tau,
                             z2* R) {
    const int lane = threadIdx.x & 63;
    double part = 0.0;
    for (int i = lane + 2; i < L; i += 64) part += col[i].x * col[i].x + col[i].y * col[i].y;
    const double tail = wave_sum(part);
    const z2 x0 = col[1];
    double t = 0.0;
    if (tail == 0.0) {
        if (writer && lane == 0) beta[j + 1] = x0;
    } else {
        const double a0 = hypot(x0.x, x0.y);
        const double xnorm = sqrt(tail + a0 * a0);
        const z2 ph = a0 == 0.0 ? z2{1.0, 0.0} : z2{x0.x / a0, x0.y / a0};
        const z2 b = {-ph.x * xnorm, -ph.y * xnorm};
        const z2 u0 = {x0.x - b.x, x0.y - b.y};
        const double iu2 = 1.0 / (u0.x * u0.x + u0.y * u0.y);
        t = 2.0 / (1.0 + tail * iu2);
        for (int i = lane + 1; i < L; i += 64) {
            const z2 a = col[i];
            const z2 w = i == 1 ? z2{1.0, 0.0} : z2{(a.x * u0.x + a.y * u0.y) * iu2, (a.y * u0.x - a.x * u0.y) * iu2};
            vout[i - 1] = w;
            if (writer && i >= 2) R[(j + i) + static_cast<long>(j) * n] = w;
        }
        if (writer && lane == 0) beta[j + 1] = b;
    }
    if (writer && lane == 0) tau[j] = t;
    return t;
}

__global__ __launch_bounds__(kThreads) void tridiag_coop(const z2* __restrict__ A, int n, z2* R, double* d,
                                                         z2* beta, double* tau, z2* phase, double* eabs,
                                                         z2* xP, z2* xC, unsigned* bar, int* fail) {
    extern __shared__ z2 sh[];
    const int W = gridDim.x, wg = blockIdx.x;
    const int ncl = (n + W - 1) / W;
    z2* vb0 = sh;
    z2* vb1 = sh + n;
    z2* q = sh + 2 * n;
    z2* p = sh + 3 * n;
    z2* colb = sh + 4 * n;
    double* red = reinterpret_cast<double*>(sh + 5 * n);  // 32 doubles
    z2* Mloc = sh + 5 * n + 16;                           // ncl columns of n
    const int tid = threadIdx.x, nt = blockDim.x, lane = tid & 63;
    const int wv = __builtin_amdgcn_readfirstlane(tid >> 6), nw = nt >> 6;
    const bool writer = wg == 0;
    int slot = 0;
    unsigned nbar = 0;
    // own columns, Hermitian part (A + A^H) / 2 with a real diagonal
    for (long e = tid; e < static_cast<long>(ncl) * n; e += nt) {
        const int c = static_cast<int>(e / n), i = static_cast<int>(e - static_cast<long>(c) * n);
        const int j = wg + W * c;
        if (j >= n) continue;
        const z2 a = A[i + static_cast<long>(j) * n], b = A[j + static_cast<long>(i) * n];
        Mloc[e] = i == j ? z2{a.x, 0.0} : z2{0.5 * (a.x + b.x), 0.5 * (a.y - b.y)};
    }
    // column 0 for the first reflector
    for (int i = tid; i < n; i += nt) {
        const z2 a = A[i], b = A[static_cast<long>(i) * n];
        colb[i] = i == 0 ? z2{a.x, 0.0} : z2{0.5 * (a.x + b.x), 0.5 * (a.y - b.y)};
    }
    if (writer && tid == 0) {
        beta[0] = {0.0, 0.0};
        for (int k = 0; k < n; ++k) tau[k] = 0.0;
    }
    __syncthreads();
    if (writer && tid == 0) d[0] = colb[0].x;
    double tk = 0.0;
    double tn = block_house(colb, n, n, 0, vb0, red, slot, writer, beta, tau, R);
    z2* v = vb1;
    z2* vn = vb0;
    for (int k = -1; k <= n - 3; ++k) {
        const int len = n - k - 2, base = k + 2, par = (k + 1) & 1;
        z2* P = xP + static_cast<long>(par) * n;
        z2* Cc = xC + static_cast<long>(par) * n;
        // own columns j >= base: local c with wg + W c >= base
        const int c_first = base <= wg ? 0 : (base - wg + W - 1) / W;
        for (int c = c_first + wv; c < ncl; c += nw) {
            const int j = wg + W * c;
            if (j >= n) break;
            z2* col = Mloc + static_cast<long>(c) * n + base;
            const int jb = j - base;
            const z2 qj = tk != 0.0 ? q[jb + 1] : z2{0.0, 0.0};
            const z2 vj = tk != 0.0 ? v[jb + 1] : z2{0.0, 0.0};
            double sr = 0.0, si = 0.0;
            for (int r = lane; r < len; r += 64) {
                z2 a = col[r];
                if (tk != 0.0) {
                    const z2 t1 = zmulc(v[r + 1], qj), t2 = zmulc(q[r + 1], vj);
                    a.x -= t1.x + t2.x;
                    a.y -= t1.y + t2.y;
                    col[r] = a;
                }
                if (jb == 0) Cc[r] = a;
                const z2 w = vn[r];
                sr += a.x * w.x + a.y * w.y;
                si += a.x * w.y - a.y * w.x;
            }
            sr = wave_sum(sr);
            si = wave_sum(si);
            if (lane == 0) P[jb] = {tn * sr, tn * si};
        }
        nbar += 1u;
        grid_barrier(bar + 32, W, nbar, fail);
        // one wavefront: p and the pivot column, f = tn vn^H p / 2, q = p - f vn, the pivot column
        // with reflector k+1, the next reflector; the other wavefronts wait at the barrier below
        if (wv == 0) {
            // p and the pivot column from the other workgroups: four rounds of loads in flight
            // before their LDS stores (a loop of load pair, wait, store waited out every round)
            for (int i0 = lane; i0 < len; i0 += 4 * 64) {
                z2 pa[4], ca[4];
#pragma unroll
                for (int u = 0; u < 4; ++u) {  // unconditional (clamped) loads: all eight in flight
                    const int i = min(i0 + 64 * u, len - 1);
                    pa[u] = P[i];
                    ca[u] = Cc[i];
                }
#pragma unroll
                for (int u = 0; u < 4; ++u) {  // a clamped index stores entry len - 1's own value again
                    const int i = min(i0 + 64 * u, len - 1);
                    p[i] = pa[u];
                    colb[i] = ca[u];
                }
            }
            double f = 0.0;
            if (tn != 0.0) {
                double part = 0.0;
                for (int i = lane; i < len; i += 64) part += vn[i].x * p[i].x + vn[i].y * p[i].y;
                f = 0.5 * tn * wave_sum(part);
                for (int i = lane; i < len; i += 64) q[i] = {p[i].x - f * vn[i].x, p[i].y - f * vn[i].y};
                wave_fence();
                const z2 q0 = q[0], v0 = vn[0];
                for (int i = lane; i < len; i += 64) {  // pivot column with reflector k+1
                    const z2 t1 = zmulc(vn[i], q0), t2 = zmulc(q[i], v0);
                    colb[i].x -= t1.x + t2.x;
                    colb[i].y -= t1.y + t2.y;
                }
            }
            wave_fence();
            if (writer && lane == 0) d[base] = colb[0].x;
            double tj = 0.0;
            if (base <= n - 3) {
                tj = wave_house(colb, len, n, base, v, writer, beta, tau, R);
            } else if (base == n - 2) {
                if (writer && lane == 0) beta[n - 1] = colb[1];
            }
            if (lane == 0) red[0] = tj;
        }
        __syncthreads();
        const double tj = red[0];
        tk = tn;
        tn = tj;
        z2* t = v;
        v = vn;
        vn = t;
        __syncthreads();
    }
    if (writer && tid == 0) {
        z2 ph = {1.0, 0.0};
        phase[0] = ph;
        eabs[0] = 0.0;
        for (int i = 1; i < n; ++i) {
            const z2 b = beta[i];
            const double r = hypot(b.x, b.y);
            eabs[i] = r;
            if (r != 0.0) ph = zmul(ph, z2{b.x / r, b.y / r});
            phase[i] = ph;
        }
    }
}

__global__ void zero_words(unsigned* a, int n) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) a[i] = 0u;
}

// The give-up flag of a multi-workgroup launch set before it starts (test hook): every grid
// barrier then returns at once, the kernel runs to its end on stale data and the host redoes
// the solve on one workgroup.
__global__ void set_word(int* a, int v) {
    if (blockIdx.x == 0 && threadIdx.x == 0) a[0] = v;
}

// ---------------------------------------------------------------- 4. eigenvalues ----
// Number of eigenvalues of the tridiagonal (d, e2 = e^2) below x (Sturm count of the LDL^T
// pivots; pivots smaller than pivmin are replaced by -pivmin, as LAPACK's dstebz does).
__device__ int sturm_count(int n, const double* d, const double* e2, double x, double pivmin) {
    int c = 0;
    double q = d[0] - x;
    if (fabs(q) <= pivmin) q = -pivmin;
    c += q < 0.0;
    for (int i = 1; i < n; ++i) {
        q = (d[i] - x) - div_ulp(e2[i], q);
        if (fabs(q) <= pivmin) q = -pivmin;
        c += q < 0.0;
    }
    return c;
}

// Gershgorin interval [gl, gu] of the tridiagonal, its norm bound tn and pivmin; lane-parallel
// over one wavefront, same result in every lane.
__device__ void tri_bounds(int n, const double* d, const double* e, double& gl, double& gu, double& tn,
                           double& pivmin) {
    const int lane = threadIdx.x & 63;
    double lo = INFINITY, hi = -INFINITY, e2max = 0.0;
    for (int i = lane; i < n; i += 64) {
        const double r = (i > 0 ? fabs(e[i]) : 0.0) + (i + 1 < n ? fabs(e[i + 1]) : 0.0);
        lo = fmin(lo, d[i] - r);
        hi = fmax(hi, d[i] + r);
        if (i > 0) e2max = fmax(e2max, e[i] * e[i]);
    }
    lo = wave_min(lo);
    hi = wave_max(hi);
    e2max = wave_max(e2max);
    tn = fmax(fabs(lo), fabs(hi));
    pivmin = DBL_MIN * fmax(1.0, e2max);
    const double pad = (n + 2) * kEps * tn + 4.0 * pivmin;
    gl = lo - pad;
    gu = hi + pad;
}

// Eigenvalue number blockIdx.x (ascending) of the tridiagonal by multisection.  One
// wavefront per eigenvalue; each round evaluates 63 interior points of the bracket.
__global__ __launch_bounds__(64) void bisect_kernel(int n, const double* __restrict__ d, const double* __restrict__ e,
                                                    double* __restrict__ w) {
    extern __shared__ double tsh[];
    double* sd = tsh;
    double* se2 = tsh + n;
    const int lane = threadIdx.x;
    for (int i = lane; i < n; i += 64) {
        sd[i] = d[i];
        se2[i] = i > 0 ? e[i] * e[i] : 0.0;
    }
    __syncthreads();
    double gl, gu, tn, pivmin;
    tri_bounds(n, d, e, gl, gu, tn, pivmin);
    const int target = blockIdx.x + 1;  // lambda_i is the smallest x with count(x) >= i + 1
    double lo = gl, hi = gu;
    for (int round = 0; round < 32; ++round) {
        if (hi - lo <= 2.0 * kEps * fmax(fabs(lo), fabs(hi)) + kEps * tn) break;
        const double x = lane == 63 ? hi : lo + (hi - lo) * static_cast<double>(lane + 1) * (1.0 / 64.0);
        const int cnt = lane == 63 ? n : sturm_count(n, sd, se2, x, pivmin);
        const unsigned long long mask = __ballot(cnt >= target);
        const int f = __ffsll(static_cast<long long>(mask)) - 1;  // lane 63 is always set
        const double nlo = f == 0 ? lo : lo + (hi - lo) * static_cast<double>(f) * (1.0 / 64.0);
        const double nhi = f == 63 ? hi : lo + (hi - lo) * static_cast<double>(f + 1) * (1.0 / 64.0);
        if (nlo == lo && nhi == hi) break;
        lo = nlo;
        hi = nhi;
    }
    if (lane == 0) w[blockIdx.x] = 0.5 * (lo + hi);
}

// ---------------------------------------------------------------- 5. eigenvectors ----
// Inverse iteration on (Tr - mu I) for the nvec computed eigenvalues; one thread per vector.
// Per-vector arrays live in global memory as [i][c] so that neighbouring threads touch
// neighbouring addresses.  Eigenvalues closer than 1e-6 |T| form clusters; the r-th member of
// a cluster is computed in round r and orthogonalised against the earlier members (the
// procedure of LAPACK's dstein with a tighter cluster criterion; the looser orthogonality of
// vectors in different clusters, about eps |T| / gap, is removed by the final Loewdin step).
// Loads of one thread's own column, kPf at a time ahead of the recurrence that consumes them:
// the recurrences are serial chains, so without this every step waits for an L2 round trip.
constexpr int kPf = 16;

constexpr int kInvitThreads = 256;  // one thread per eigenvector; registers for the prefetch buffers

__global__ __launch_bounds__(kInvitThreads) void invit_kernel(int n, const double* __restrict__ dg,
                                                         const double* __restrict__ eg, double* w, int nvec,
                                                         double* __restrict__ Zt, double* __restrict__ U1,
                                                         double* __restrict__ U2, double* __restrict__ U3,
                                                         double* __restrict__ LO, int* __restrict__ SW) {
    // d, e of the tridiagonal and the eigenvalues w in LDS: every thread reads them at every
    // step, and the ordering pass below is a serial chain
    extern __shared__ double tde[];
    double* d = tde;
    double* e = tde + n;
    double* ws = tde + 2 * n;
    __shared__ double red[32];
    __shared__ double sgl[4];
    int slot = 0;
    const int tid = threadIdx.x, nt = blockDim.x;
    for (int i = tid; i < n; i += nt) {
        d[i] = dg[i];
        e[i] = eg[i];
    }
    for (int c = tid; c < nvec; c += nt) ws[c] = w[c];
    __syncthreads();
    if (tid < 64) {
        double gl, gu, tn, pivmin;
        tri_bounds(n, d, e, gl, gu, tn, pivmin);
        if (tid == 0) sgl[0] = tn;
    }
    if (tid == 0)  // eigenvalues from independent bisections: enforce ascending order
        for (int c = 1; c < nvec; ++c) ws[c] = fmax(ws[c], ws[c - 1]);
    __syncthreads();
    for (int c = tid; c < nvec; c += nt) w[c] = ws[c];
    const double tn = sgl[0];
    // Levels closer than ctol form a cluster (LU inverse iteration with Gram-Schmidt inside the
    // cluster).  Above it the twisted vector is off by about eps tn / gap <= 2e-8, and the
    // back-transformation's Y (1.5 I - 0.5 Y^H Y) step leaves its square, ~5e-16.
    const double ctol = 1e-8 * tn;
    const double pivmin = fmax(kEps * tn, 1e-300);
    int maxrank = 0;
    for (int c = tid; c < nvec; c += nt) {
        int c0 = c;
        while (c0 > 0 && ws[c0] - ws[c0 - 1] <= ctol) --c0;
        maxrank = max(maxrank, c - c0);
    }
    const int rounds = static_cast<int>(block_max(static_cast<double>(maxrank), red, slot)) + 1;
    for (int r = 0; r < rounds; ++r) {
        for (int c = tid; c < nvec; c += nt) {
            int c0 = c;
            while (c0 > 0 && ws[c0] - ws[c0 - 1] <= ctol) --c0;
            if (c - c0 != r) continue;
            // shifts of a cluster kept apart by at least 10 eps |mu|
            double mu = ws[c0];
            for (int cc = c0 + 1; cc <= c; ++cc) {
                double m = ws[cc];
                if (m - mu < 10.0 * kEps * fabs(m)) m = mu + 10.0 * kEps * fmax(fabs(m), tn * kEps);
                mu = m;
            }
            auto at = [nvec, c](int i) { return static_cast<long>(i) * nvec + c; };
            const bool isolated = c == c0 && (c + 1 >= nvec || ws[c + 1] - ws[c] > ctol);
            if (isolated && n >= 2) {
                // Twisted factorisation of Tr - mu I (the core step of the MRRR algorithm,
                // Dhillon & Parlett, LAPACK's dlar1v): L+ D+ L+^T from the top, U- D- U-^T from
                // the bottom, twist index r where |gamma_r| = |D+_r + D-_r - (d_r - mu)| is
                // smallest, then z_r = 1 and the two one-term recurrences outwards.  One pass
                // each way instead of repeated LU solves; for an eigenvalue apart from the others
                // the vector is accurate to about eps |T| / gap.  U1 holds L+, U2 D+, U3 U-.
                // The divisions are div_ulp's: the recurrences need O(eps) relative error per
                // operation, and they are the serial chain of this kernel.
                double dp = d[0] - mu;
                for (int i = 0; i + 1 < n; ++i) {
                    if (fabs(dp) < pivmin) dp = dp >= 0.0 ? pivmin : -pivmin;
                    const double l = div_ulp(e[i + 1], dp);
                    U1[at(i)] = l;
                    U2[at(i)] = dp;
                    dp = (d[i + 1] - mu) - l * e[i + 1];
                }
                if (fabs(dp) < pivmin) dp = dp >= 0.0 ? pivmin : -pivmin;
                U2[at(n - 1)] = dp;
                double dm = d[n - 1] - mu;
                double best = fabs(dp + dm - (d[n - 1] - mu));
                int rt = n - 1;
                for (int i0 = n - 1; i0 >= 1; i0 -= kPf) {  // steps i = i0, i0-1, ... >= 1
                    double dpb[kPf];
#pragma unroll
                    for (int k = 0; k < kPf; ++k)
                        if (i0 - k >= 1) dpb[k] = U2[at(i0 - k - 1)];
#pragma unroll
                    for (int k = 0; k < kPf; ++k) {
                        const int i = i0 - k;
                        if (i < 1) break;
                        if (fabs(dm) < pivmin) dm = dm >= 0.0 ? pivmin : -pivmin;
                        const double u = div_ulp(e[i], dm);
                        U3[at(i)] = u;
                        dm = (d[i - 1] - mu) - u * e[i];
                        const double g = fabs(dpb[k] + dm - (d[i - 1] - mu));
                        if (g < best) {
                            best = g;
                            rt = i - 1;
                        }
                    }
                }
                double z = 1.0, nrm = 1.0;
                Zt[at(rt)] = 1.0;
                for (int i0 = rt - 1; i0 >= 0; i0 -= kPf) {
                    double lb[kPf];
#pragma unroll
                    for (int k = 0; k < kPf; ++k)
                        if (i0 - k >= 0) lb[k] = U1[at(i0 - k)];
#pragma unroll
                    for (int k = 0; k < kPf; ++k) {
                        const int i = i0 - k;
                        if (i < 0) break;
                        z = -lb[k] * z;
                        Zt[at(i)] = z;
                        nrm += z * z;
                    }
                }
                z = 1.0;
                for (int i0 = rt + 1; i0 < n; i0 += kPf) {
                    double ub[kPf];
#pragma unroll
                    for (int k = 0; k < kPf; ++k)
                        if (i0 + k < n) ub[k] = U3[at(i0 + k)];
#pragma unroll
                    for (int k = 0; k < kPf; ++k) {
                        const int i = i0 + k;
                        if (i >= n) break;
                        z = -ub[k] * z;
                        Zt[at(i)] = z;
                        nrm += z * z;
                    }
                }
                const double inv = 1.0 / sqrt(nrm);
                for (int i0 = 0; i0 < n; i0 += kPf) {
                    double zb[kPf];
#pragma unroll
                    for (int k = 0; k < kPf; ++k)
                        if (i0 + k < n) zb[k] = Zt[at(i0 + k)];
#pragma unroll
                    for (int k = 0; k < kPf; ++k)
                        if (i0 + k < n) Zt[at(i0 + k)] = zb[k] * inv;
                }
                continue;
            }
            // LU with partial pivoting of Tr - mu I: row i keeps (u1 diag, u2, u3 super-diagonals).
            // The pending row i+1 is carried in registers, so the loop only stores.
            double u1 = d[0] - mu, u2 = n > 1 ? e[1] : 0.0;
#pragma unroll 4
            for (int i = 0; i + 1 < n; ++i) {
                const double sub = e[i + 1];
                const double b2 = d[i + 1] - mu, b3 = i + 2 < n ? e[i + 2] : 0.0;  // row i+1 as it stands
                if (fabs(u1) >= fabs(sub)) {
                    if (fabs(u1) < pivmin) u1 = copysign(pivmin, u1 == 0.0 ? 1.0 : u1);
                    const double f = sub / u1;
                    SW[at(i)] = 0;
                    LO[at(i)] = f;
                    U1[at(i)] = u1;
                    U2[at(i)] = u2;
                    U3[at(i)] = 0.0;
                    u1 = b2 - f * u2;
                    u2 = b3;
                } else {
                    const double f = u1 / sub;
                    SW[at(i)] = 1;
                    LO[at(i)] = f;
                    U1[at(i)] = sub;
                    U2[at(i)] = b2;
                    U3[at(i)] = b3;
                    u1 = u2 - f * b2;
                    u2 = -f * b3;
                }
            }
            if (fabs(u1) < pivmin) u1 = copysign(pivmin, u1 == 0.0 ? 1.0 : u1);
            U1[at(n - 1)] = u1;
            U2[at(n - 1)] = 0.0;
            U3[at(n - 1)] = 0.0;
            for (int i = 0; i < n; ++i) Zt[at(i)] = 1.0 + 0.1 * sin(1.3 * i + 0.7 * c);
            for (int it = 0; it < 3; ++it) {
                // forward: L^-1 with the recorded interchanges
                double xi = Zt[at(0)];
#pragma unroll 8
                for (int i = 0; i + 1 < n; ++i) {
                    double xn = Zt[at(i + 1)];
                    if (SW[at(i)]) {
                        const double t = xi;
                        xi = xn;
                        xn = t;
                    }
                    Zt[at(i)] = xi;
                    xi = xn - LO[at(i)] * xi;
                }
                Zt[at(n - 1)] = xi;
                // backward: U x = y
                double x1 = 0.0, x2 = 0.0;  // x[i+1], x[i+2]
#pragma unroll 8
                for (int i = n - 1; i >= 0; --i) {
                    const double s = Zt[at(i)] - U2[at(i)] * x1 - U3[at(i)] * x2;
                    const double xv = s / U1[at(i)];
                    Zt[at(i)] = xv;
                    x2 = x1;
                    x1 = xv;
                }
                // against the earlier members of the cluster
                for (int qv = c0; qv < c; ++qv) {
                    double dot = 0.0;
                    for (int i = 0; i < n; ++i) dot += Zt[static_cast<long>(i) * nvec + qv] * Zt[at(i)];
                    for (int i = 0; i < n; ++i) Zt[at(i)] -= dot * Zt[static_cast<long>(i) * nvec + qv];
                }
                double nrm = 0.0;
                for (int i = 0; i < n; ++i) nrm += Zt[at(i)] * Zt[at(i)];
                const double inv = nrm > 0.0 ? 1.0 / sqrt(nrm) : 0.0;
                for (int i = 0; i < n; ++i) Zt[at(i)] *= inv;
            }
        }
        __syncthreads();
    }
}

// ---------------------------------------------------------------- 6. back-transformation ----
// Y[:, c] = H_0 H_1 ... H_{n-3} D Zt[:, c]; one wavefront per column, RPL rows per lane.
template <int RPL>
__global__ __launch_bounds__(64) void backtransform_kernel(int n, const z2* __restrict__ A,
                                                           const double* __restrict__ tau,
                                                           const z2* __restrict__ phase,
                                                           const double* __restrict__ Zt, int nvec,
                                                           z2* __restrict__ Y) {
    const int lane = threadIdx.x, c = blockIdx.x;
    z2 y[RPL];
#pragma unroll
    for (int r = 0; r < RPL; ++r) {
        const int i = lane + 64 * r;
        if (i < n) {
            const double z = Zt[static_cast<long>(i) * nvec + c];
            const z2 ph = phase[i];
            y[r] = {ph.x * z, ph.y * z};
        } else {
            y[r] = {0.0, 0.0};
        }
    }
    // reflectors k = n-3 .. 0, PF at a time: their entries are loaded together, then applied in
    // order, so one L2 round trip serves PF serial steps
    constexpr int PF = RPL <= 16 ? 32 / RPL : 2;
    for (int k0 = n - 3; k0 >= 0; k0 -= PF) {
        z2 vb[PF][RPL];
        double tb[PF];
#pragma unroll
        for (int q = 0; q < PF; ++q) {
            const int k = k0 - q;
            tb[q] = k >= 0 ? tau[k] : 0.0;
            const z2* col = A + static_cast<long>(k >= 0 ? k : 0) * n;
#pragma unroll
            for (int r = 0; r < RPL; ++r) {
                const int i = lane + 64 * r;
                vb[q][r] = (k >= 0 && i > k + 1 && i < n) ? col[i] : z2{i == k + 1 ? 1.0 : 0.0, 0.0};
            }
        }
#pragma unroll
        for (int q = 0; q < PF; ++q) {
            const double t = tb[q];
            if (t == 0.0) continue;  // also every k < 0 of the last chunk
            double sr = 0.0, si = 0.0;
#pragma unroll
            for (int r = 0; r < RPL; ++r) {
                sr += vb[q][r].x * y[r].x + vb[q][r].y * y[r].y;
                si += vb[q][r].x * y[r].y - vb[q][r].y * y[r].x;
            }
            sr = t * wave_sum(sr);
            si = t * wave_sum(si);
#pragma unroll
            for (int r = 0; r < RPL; ++r) {
                y[r].x -= vb[q][r].x * sr - vb[q][r].y * si;
                y[r].y -= vb[q][r].x * si + vb[q][r].y * sr;
            }
        }
    }
#pragma unroll
    for (int r = 0; r < RPL; ++r) {
        const int i = lane + 64 * r;
        if (i < n) Y[i + static_cast<long>(c) * n] = y[r];
    }
}

// G <- 1.5 I - 0.5 G (m x m)
__global__ void ortho_factor(int m, z2* __restrict__ G) {
    const long mm = static_cast<long>(m) * m;
    for (long e = blockIdx.x * static_cast<long>(blockDim.x) + threadIdx.x; e < mm;
         e += static_cast<long>(gridDim.x) * blockDim.x) {
        const z2 g = G[e];
        G[e] = {((e % m) == (e / m) ? 1.5 : 0.0) - 0.5 * g.x, -0.5 * g.y};
    }
}

struct Layout {
    size_t off[32];  // >= B_COUNT
    size_t total;
};
enum Buf { B_HF, B_SF, B_ZL, B_C, B_T1, B_Y, B_Z2, B_ZY, B_YB, B_G, B_Y2, B_D, B_E, B_TAU, B_W, B_ZT, B_U1, B_U2,
           B_U3, B_LO, B_SW, B_BETA, B_PH, B_ERR, B_XP, B_XC, B_BAR, B_COUNT };
static_assert(B_COUNT <= 32, "Layout::off holds 32 buffers");
Layout layout(int n, int nvec) {
    const size_t nn = static_cast<size_t>(n) * n * 16, nv = static_cast<size_t>(n) * nvec;
    const size_t sz[B_COUNT] = {nn, nn, nn, nn, nn, nn, nn, nn,                   // Hf Sf Zl C T1 Y Z2 ZY
                                nv * 16, static_cast<size_t>(nvec) * nvec * 16, nv * 16,  // Yb G Y2
                                size_t(n) * 8, size_t(n) * 8, size_t(n) * 8, size_t(nvec) * 8,  // d e tau w
                                nv * 8, nv * 8, nv * 8, nv * 8, nv * 8, nv * 4,  // Zt U1 U2 U3 LO SW
                                size_t(n) * 16, size_t(n) * 16, 64,               // beta phase err
                                size_t(n) * 32, size_t(n) * 32, (size_t(n) + 1) * 128};  // exchange p, pivot; give-up word, W <= n flags
    Layout L{};
    size_t o = 0;
    for (int i = 0; i < B_COUNT; ++i) {
        L.off[i] = o;
        o += (sz[i] + 255) & ~size_t(255);
    }
    L.total = o;
    return L;
}

size_t tridiag_lds(int n, bool lm) { return (6 * static_cast<size_t>(n) + 17 + (lm ? static_cast<size_t>(n) * n : 0)) * 16; }

// multi-workgroup launches still to be started with their give-up flag set (test hook), and
// redos on one workgroup since the process started
std::atomic<int> g_forced_giveups{0};
std::atomic<long> g_barrier_redos{0};
// eigensolves to fail on purpose (test hook): bit 0 is the next solve; shifted once per solve
std::atomic<unsigned long long> g_fail_mask{0};

bool take_forced_failure() {
    unsigned long long m = g_fail_mask.load();
    while (m != 0)
        if (g_fail_mask.compare_exchange_weak(m, m >> 1)) return (m & 1ull) != 0;
    return false;
}

bool take_forced_giveup() {
    int v = g_forced_giveups.load();
    while (v > 0)
        if (g_forced_giveups.compare_exchange_weak(v, v - 1)) return true;
    return false;
}

}  // namespace

void force_barrier_giveups(int count) { g_forced_giveups.store(std::max(0, count)); }
void fail_eigensolves(unsigned long long mask) { g_fail_mask.store(mask); }
long barrier_redos_total() { return g_barrier_redos.load(); }

size_t subspace_eig_workspace(int n, int nvec) { return layout(std::max(n, 1), std::max(nvec, 1)).total; }

int subspace_eig(int n, const double* H, long ldh, const double* S, long lds, int nvec, double* w_host, double* X,
                 long ldx, void* work, size_t work_bytes, void* gemm_work, size_t gemm_bytes, size_t max_lds,
                 int num_cus_hint, hipStream_t st, hipError_t* err, SubspaceEigStats* stats, Profiler* prof, HostStage* pin,
                 double lowdin_bound) {
    // stage timing on the caller's profiler (events on this stream; nothing when it is off)
    int mark_prev = prof ? prof->mark(st) : -1;
    auto stage_done = [&](int id) {
        if (!prof || !prof->on) return;
        const int m = prof->mark(st);
        prof->span(id, mark_prev, m);
        mark_prev = m;
    };
    *err = hipSuccess;
    if (n <= 0 || nvec <= 0) return 0;
    if (take_forced_failure()) return n + 1;  // test hook: reported as a non-positive-definite S
    nvec = std::min(nvec, n);
    const Layout L = layout(n, nvec);
    if (work_bytes < L.total) {
        *err = hipErrorInvalidValue;
        return -1;
    }
    char* base = static_cast<char*>(work);
    auto P = [&](int b) { return reinterpret_cast<z2*>(base + L.off[b]); };
    auto D = [&](int b) { return reinterpret_cast<double*>(base + L.off[b]); };
    z2 *Hf = P(B_HF), *Sf = P(B_SF), *Zl = P(B_ZL), *C = P(B_C), *T1 = P(B_T1);
    double* errd = D(B_ERR);
    auto ok = [&](hipError_t e) {
        if (e != hipSuccess && *err == hipSuccess) *err = e;
        return e == hipSuccess;
    };
    auto gemm = [&](int opa, int m, int nc, int k, const void* A, long lda, const void* B, long ldb, void* Cm,
                    long ldc) {
        return ok(zgemm(GemmTypes::D_D_D_D, opa, m, nc, k, 1.0, A, lda, B, ldb, 0.0, Cm, ldc, gemm_work, gemm_bytes,
                        st));
    };
    const long nn = static_cast<long>(n) * n;
    const unsigned eg = static_cast<unsigned>(std::min<long>((nn + 255) / 256, 1024));

    hipLaunchKernelGGL(zero_words, dim3(1), dim3(64), 0, st, reinterpret_cast<unsigned*>(errd), 2);
    hipLaunchKernelGGL(herm_prep, dim3(static_cast<unsigned>(std::min<long>((nn + 255) / 256, 512))), dim3(256), 0,
                       st, n, reinterpret_cast<const z2*>(H), ldh, reinterpret_cast<const z2*>(S), lds, Hf, Sf, Zl,
                       reinterpret_cast<unsigned long long*>(errd));
    if (!ok(hipGetLastError())) return -1;
    stage_done(K_EIG_PREP);
    // Newton-Schulz on S / c with c >= lambda_max(S): Y -> (S/c)^(1/2), Z -> (S/c)^(-1/2);
    // replaces the first-order factor in Zl.  0, n + 1 (no convergence: S not positive
    // definite) or -1.
    auto newton_schulz = [&](double sdev) -> int {
        if (!(sdev == sdev) || sdev > 1e6) return n + 1;
        const double c = 1.0 + n * sdev;
        z2 *Y = P(B_Y), *Z = Zl, *Z2 = P(B_Z2), *ZY = P(B_ZY), *Tm = T1, *Yn = C;
        hipLaunchKernelGGL(ns_init, dim3(eg), dim3(256), 0, st, n, Sf, 1.0 / c, Y, Z);
        if (!ok(hipGetLastError())) return -1;
        bool conv = false;
        int it = 0;
        for (; it < 80 && !conv; ++it) {
            if (!gemm(0, n, n, n, Z, n, Y, n, ZY, n)) return -1;
            hipLaunchKernelGGL(ns_residual, dim3(1), dim3(kThreads), 0, st, n, ZY, Tm, errd);
            double r = 0.0;
            if (!ok(hipGetLastError())) return -1;
            if (pin) {
                if (!ok(pin->d2h(&r, errd, 8, st))) return -1;
            } else if (!ok(hipMemcpyAsync(&r, errd, 8, hipMemcpyDeviceToHost, st)) || !ok(hipStreamSynchronize(st))) {
                return -1;
            }
            if (!(r < 1e3)) return n + 1;  // diverging: S is not positive definite
            conv = r < 1e-8;                // the update below squares the error
            if (!gemm(0, n, n, n, Y, n, Tm, n, Yn, n) || !gemm(0, n, n, n, Tm, n, Z, n, Z2, n)) return -1;
            std::swap(Y, Yn);
            std::swap(Z, Z2);
        }
        if (!conv) return n + 1;
        if (Z != Zl && !ok(hipMemcpyAsync(Zl, Z, nn * 16, hipMemcpyDeviceToDevice, st))) return -1;
        hipLaunchKernelGGL(scale_matrix, dim3(eg), dim3(256), 0, st, nn, 1.0 / std::sqrt(c), Zl);
        if (!ok(hipGetLastError())) return -1;
        if (stats) stats->newton_schulz_iterations = it;
        return 0;
    };
    // C = Zl Hf Zl, then steps 3-6; one synchronisation at the end, which also brings back
    // max |S - I| (errd, written by herm_prep) into *sdev_out.  0 or -1.
    // The multi-workgroup tridiagonalisation is an ordinary launch of at most one workgroup per
    // CU; if another stream's work keeps one of its workgroups from being placed until a grid
    // barrier gives up, the solve is redone with the one-workgroup kernel (return code -2).
    bool multi_wg_ok = true;
    auto solve_from_zl = [&](double* sdev_out) -> int {
        // C = Zl Hf Zl
        if (!gemm(0, n, n, n, Hf, n, Zl, n, T1, n) || !gemm(0, n, n, n, Zl, n, T1, n, C, n)) return -1;
        stage_done(K_EIG_GEMM);
        // tridiagonalisation
        double *d = D(B_D), *e = D(B_E), *tau = D(B_TAU), *w = D(B_W);
        z2 *beta = P(B_BETA), *ph = P(B_PH);
        const int rows_lane = (n + 63) / 64;
        const bool lm = rows_lane <= 2 && tridiag_lds(n, true) <= max_lds;
        const size_t tl = tridiag_lds(n, lm);
        if (tl > max_lds || rows_lane > 16) {
            *err = hipErrorInvalidValue;
            return -1;
        }
        static std::atomic<bool> lds_raised{false};  // set once per process; the attribute call is idempotent
        if (!lds_raised && max_lds > 65536) {
            if (!ok(hipFuncSetAttribute(reinterpret_cast<const void*>(&tridiag_kernel<true, 1>),
                                        hipFuncAttributeMaxDynamicSharedMemorySize, static_cast<int>(max_lds))) ||
                !ok(hipFuncSetAttribute(reinterpret_cast<const void*>(&tridiag_kernel<true, 2>),
                                        hipFuncAttributeMaxDynamicSharedMemorySize, static_cast<int>(max_lds))))
                return -1;
            lds_raised = true;
        }
            // several CUs when the matrix does not fit one CU's LDS: column j on workgroup j mod W,
        // about 16 columns (one per wavefront) each, fewer when LDS is short
        z2* refl = C;  // where the reflectors end up
        bool coop_done = false;
        if (!lm && n >= 100 && multi_wg_ok) {
            const long room = static_cast<long>(max_lds / 16) - 5L * n - 16;
            const int ncl = static_cast<int>(std::min<long>(16, room / n));
            if (ncl >= 1) {
                const int W = (n + ncl - 1) / ncl;
                const size_t lds = (5 * static_cast<size_t>(n) + 16 + static_cast<size_t>((n + W - 1) / W) * n) * 16;
                if (W <= num_cus_hint && lds <= max_lds) {
                    static std::atomic<bool> coop_raised{false};
                    if (!coop_raised) {
                        if (!ok(hipFuncSetAttribute(reinterpret_cast<const void*>(&tridiag_coop),
                                                    hipFuncAttributeMaxDynamicSharedMemorySize,
                                                    static_cast<int>(max_lds))))
                            return -1;
                        coop_raised = true;
                    }
                    // word 1: the give-up flag; word 32 (w + 1): workgroup w's step flag (W <= n)
                    unsigned* bar = reinterpret_cast<unsigned*>(base + L.off[B_BAR]);
                    hipLaunchKernelGGL(zero_words, dim3((32 + 32 * W + 63) / 64), dim3(64), 0, st, bar, 32 + 32 * W);
                    if (!ok(hipGetLastError())) return -1;
                    const z2* Cc = C;
                    z2* R = T1;
                    z2* xp = P(B_XP);
                    z2* xc = P(B_XC);
                    int* fail = reinterpret_cast<int*>(bar + 1);
                    if (take_forced_giveup()) {
                        hipLaunchKernelGGL(set_word, dim3(1), dim3(64), 0, st, fail, 1);
                        if (!ok(hipGetLastError())) return -1;
                    }
                    hipLaunchKernelGGL(tridiag_coop, dim3(W), dim3(kThreads), lds, st, Cc, n, R, d, beta, tau, ph, e,
                                       xp, xc, bar, fail);
                    if (!ok(hipGetLastError())) return -1;
                    coop_done = true;
                    refl = T1;
                }
            }
        }
#define Q3X_TRD(LMV, R) \
        hipLaunchKernelGGL((tridiag_kernel<LMV, R>), dim3(1), dim3(kThreads), tl, st, C, n, d, beta, tau, ph, e)
        if (coop_done) {
        } else if (lm) {
            if (rows_lane <= 1) Q3X_TRD(true, 1);
            else Q3X_TRD(true, 2);
        } else if (rows_lane <= 1) Q3X_TRD(false, 1);
        else if (rows_lane <= 2) Q3X_TRD(false, 2);
        else if (rows_lane <= 4) Q3X_TRD(false, 4);
        else if (rows_lane <= 6) Q3X_TRD(false, 6);
        else if (rows_lane <= 8) Q3X_TRD(false, 8);
        else if (rows_lane <= 12) Q3X_TRD(false, 12);
        else Q3X_TRD(false, 16);
    #undef Q3X_TRD
        if (!ok(hipGetLastError())) return -1;
        stage_done(K_EIG_TRIDIAG);
        hipLaunchKernelGGL(bisect_kernel, dim3(nvec), dim3(64), static_cast<size_t>(n) * 16, st, n, d, e, w);
        if (!ok(hipGetLastError())) return -1;
        stage_done(K_EIG_BISECT);
        hipLaunchKernelGGL(invit_kernel, dim3(1), dim3(kInvitThreads), static_cast<size_t>(2 * n + nvec) * 8, st, n, d, e, w, nvec, D(B_ZT), D(B_U1), D(B_U2),
                           D(B_U3), D(B_LO), reinterpret_cast<int*>(base + L.off[B_SW]));
        if (!ok(hipGetLastError())) return -1;
        stage_done(K_EIG_INVIT);
        z2* Yb = P(B_YB);
        const int rpl = (n + 63) / 64;
    #define Q3X_BT(R)                                                                                              \
        hipLaunchKernelGGL(backtransform_kernel<R>, dim3(nvec), dim3(64), 0, st, n, refl, tau, ph, D(B_ZT), nvec, Yb)
        if (rpl <= 1) Q3X_BT(1);
        else if (rpl <= 2) Q3X_BT(2);
        else if (rpl <= 4) Q3X_BT(4);
        else if (rpl <= 8) Q3X_BT(8);
        else if (rpl <= 16) Q3X_BT(16);
        else {
            *err = hipErrorInvalidValue;
            return -1;
        }
    #undef Q3X_BT
        if (!ok(hipGetLastError())) return -1;
        // Y <- Y (1.5 I - 0.5 Y^H Y); X = Zl Y
        z2 *G = P(B_G), *Y2 = P(B_Y2);
        if (!gemm(1, nvec, nvec, n, Yb, n, Yb, n, G, nvec)) return -1;
        const long mm = static_cast<long>(nvec) * nvec;
        hipLaunchKernelGGL(ortho_factor, dim3(static_cast<unsigned>(std::min<long>((mm + 255) / 256, 1024))), dim3(256),
                           0, st, nvec, G);
        if (!ok(hipGetLastError())) return -1;
        if (!gemm(0, n, nvec, nvec, Yb, n, G, nvec, Y2, n) || !gemm(0, n, nvec, n, Zl, n, Y2, n, X, ldx)) return -1;
        stage_done(K_EIG_BACK);
        int coop_fail = 0;
        const size_t wb = static_cast<size_t>(nvec) * 8;
        char* sw = pin ? pin->take(wb + 16, st) : nullptr;
        if (pin && !sw) {
            *err = hipErrorOutOfMemory;
            return -1;
        }
        double* w_dst = sw ? reinterpret_cast<double*>(sw) : w_host;
        double* sdev_dst = sw ? reinterpret_cast<double*>(sw + wb) : sdev_out;
        int* fail_dst = sw ? reinterpret_cast<int*>(sw + wb + 8) : &coop_fail;
        if (!ok(hipMemcpyAsync(w_dst, w, wb, hipMemcpyDeviceToHost, st)) ||
            !ok(hipMemcpyAsync(sdev_dst, errd, 8, hipMemcpyDeviceToHost, st)) ||
            (coop_done && !ok(hipMemcpyAsync(fail_dst, base + L.off[B_BAR] + 4, 4, hipMemcpyDeviceToHost, st))) ||
            !ok(hipStreamSynchronize(st)))
            return -1;
        if (sw) {
            std::memcpy(w_host, sw, wb);
            std::memcpy(sdev_out, sw + wb, 8);
            if (coop_done) std::memcpy(&coop_fail, sw + wb + 8, 4);
            pin->rewind();
        }
        if (coop_fail) return -2;  // a grid barrier gave up waiting
        return 0;
    };
    double sdev = 0.0, sdev2 = 0.0;
    if (stats) stats->newton_schulz_iterations = 0;
    auto solve_once = [&](double* sdev_out) -> int {
        int rc = solve_from_zl(sdev_out);
        if (rc == -2) {
            g_barrier_redos.fetch_add(1);
            if (stats) ++stats->barrier_redos;
            multi_wg_ok = false;
            rc = solve_from_zl(sdev_out);
            multi_wg_ok = true;
        }
        return rc;
    };
    int r = solve_once(&sdev);
    if (r != 0) return r;
    if (stats) stats->s_deviation = sdev;
    // first-order Loewdin factor error ~ (3/8) |E|_2^2 <= (3/8) (n max|E|)^2
    if (!(sdev * n <= lowdin_bound)) {
        if ((r = newton_schulz(sdev)) != 0) return r;
        if ((r = solve_once(&sdev2)) != 0) return r;
    }
    for (int i = 0; i < nvec; ++i)
        if (!(w_host[i] == w_host[i])) return n + 1;
    return 0;
}

}  // namespace q3x
