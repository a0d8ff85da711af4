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
//   5. Eigenvectors of Tr by inverse iteration, one thread per vector, in parallel; vectors
//      of (nearly) equal eigenvalues are orthogonalised against each other in rounds.
//   6. Y = Q D Z_tr (reflectors applied per column, one wavefront each), Y <- Y (1.5 I - 0.5
//      Y^H Y) to restore orthonormality to rounding, X = Z Y.
#include "subspace_eig.h"

#include <hip/hip_runtime.h>

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

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
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
// err[0] = max |S - I| (componentwise).  One workgroup.
__global__ __launch_bounds__(kThreads) void herm_prep(int n, const z2* __restrict__ H, long ldh,
                                                      const z2* __restrict__ S, long lds, z2* __restrict__ Hf,
                                                      z2* __restrict__ Sf, z2* __restrict__ Zl, double* err) {
    __shared__ double red[32];
    int slot = 0;
    double dev = 0.0;
    const long nn = static_cast<long>(n) * n;
    for (long e = threadIdx.x; e < nn; e += blockDim.x) {
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
    const double m = block_max(dev, red, slot);
    if (threadIdx.x == 0) err[0] = m;
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
// Reflector of column k, rows [k+1, n): H = I - tau v v^H, v_0 = 1, H^H x = beta e_0.
// v goes to vout (length n-k-1) and, below its leading 1, into column k rows k+2...
// Returns tau (0: no reflector, x is already of the form beta e_0).
__device__ double house(z2* M, int n, int k, z2* vout, double* red, int& slot, z2* beta_k1, double* tau_k) {
    const int tid = threadIdx.x, nt = blockDim.x;
    const int m = n - k - 1;
    z2* x = M + (k + 1) + static_cast<long>(k) * n;
    double part = 0.0;
    for (int i = tid + 1; i < m; i += nt) part += x[i].x * x[i].x + x[i].y * x[i].y;
    const double tail = block_sum(part, red, slot);
    const z2 x0 = x[0];
    double t = 0.0;
    if (tail == 0.0) {
        if (tid == 0) *beta_k1 = x0;
    } else {
        const double a0 = hypot(x0.x, x0.y);
        const double xnorm = sqrt(tail + a0 * a0);
        const z2 ph = a0 == 0.0 ? z2{1.0, 0.0} : z2{x0.x / a0, x0.y / a0};
        const z2 b = {-ph.x * xnorm, -ph.y * xnorm};
        const z2 u0 = {x0.x - b.x, x0.y - b.y};
        const double u2 = u0.x * u0.x + u0.y * u0.y;
        const double iu2 = 1.0 / u2;
        t = 2.0 / (1.0 + tail * iu2);
        for (int i = tid; i < m; i += nt) {
            const z2 xi = x[i];
            vout[i] = i == 0 ? z2{1.0, 0.0} : z2{(xi.x * u0.x + xi.y * u0.y) * iu2, (xi.y * u0.x - xi.x * u0.y) * iu2};
        }
        if (tid == 0) *beta_k1 = b;
    }
    if (tid == 0) *tau_k = t;
    __syncthreads();  // vout complete; every read of x finished
    if (t != 0.0)
        for (int i = tid + 1; i < m; i += nt) x[i] = vout[i];
    return t;
}

// Fused pass of step k over columns j in [k+2, n), rows [k+2, n):
//   a -= v_i conj(q_j) + q_i conj(v_j)     (rank-2 update of step k; v, q index rows from k+1)
//   p[j-k-2] = tn * sum_i conj(a_i) vn_i   (product of the updated block with v_{k+1})
// Column j is owned by wavefront (j - k - 2) mod nw; lanes run over rows.
__device__ void fused_pass(z2* M, int n, int k, double tk, const z2* v, const z2* q, double tn, const z2* vn, z2* p) {
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6, nw = blockDim.x >> 6;
    const int len = n - k - 2;  // rows and columns of the pass
    for (int c = wv; c < len; c += nw) {
        z2* col = M + static_cast<long>(k + 2 + c) * n + (k + 2);
        const z2 qj = tk != 0.0 ? q[c + 1] : z2{0.0, 0.0};
        const z2 vj = tk != 0.0 ? v[c + 1] : z2{0.0, 0.0};
        double sr = 0.0, si = 0.0;
        for (int r = lane; r < len; r += 64) {
            z2 a = col[r];
            if (tk != 0.0) {
                const z2 t1 = zmulc(v[r + 1], qj), t2 = zmulc(q[r + 1], vj);
                a.x -= t1.x + t2.x;
                a.y -= t1.y + t2.y;
                col[r] = a;
            }
            if (tn != 0.0) {
                const z2 w = vn[r];
                sr += a.x * w.x + a.y * w.y;
                si += a.x * w.y - a.y * w.x;
            }
        }
        if (tn != 0.0) {
            sr = wave_sum(sr);
            si = wave_sum(si);
            if (lane == 0) p[c] = {tn * sr, tn * si};
        }
    }
    __syncthreads();
}

// q = p - (tn/2) (vn^H p) vn over len entries
__device__ void form_q(int len, double tn, const z2* vn, const z2* p, z2* q, double* red, int& slot) {
    const int tid = threadIdx.x, nt = blockDim.x;
    double part = 0.0;
    for (int i = tid; i < len; i += nt) part += vn[i].x * p[i].x + vn[i].y * p[i].y;
    const double s = block_sum(part, red, slot);
    const double f = 0.5 * tn * s;
    for (int i = tid; i < len; i += nt) q[i] = {p[i].x - f * vn[i].x, p[i].y - f * vn[i].y};
    __syncthreads();
}

// A (n x n, ld n, full storage, Hermitian up to rounding) -> d (diagonal), beta (sub-diagonal,
// beta[i] couples i-1 and i), tau (reflectors, kept in A below the sub-diagonal), the real
// tridiagonal |beta| in eabs and the phases D in phase (T = D Tr D^H).
template <bool LM>
__global__ __launch_bounds__(kThreads) void tridiag_kernel(z2* A, int n, double* d, z2* beta, double* tau,
                                                           z2* phase, double* eabs) {
    extern __shared__ z2 sh[];
    z2* vb0 = sh;
    z2* vb1 = sh + n;
    z2* q = sh + 2 * n;
    z2* p = sh + 3 * n;
    double* red = reinterpret_cast<double*>(sh + 4 * n);  // 32 doubles
    z2* M = LM ? sh + 4 * n + 16 : A;
    const int tid = threadIdx.x, nt = blockDim.x;
    const long nn = static_cast<long>(n) * n;
    int slot = 0;
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
    __syncthreads();
    if (tid == 0) {
        d[0] = M[0].x;
        beta[0] = {0.0, 0.0};
        for (int k = 0; k < n; ++k) tau[k] = 0.0;
    }
    if (n == 2 && tid == 0) {
        d[1] = M[3].x;
        beta[1] = M[1];
    }
    if (n >= 3) {
        // prologue: reflector 0, p_0 = t0 B v_0, q_0
        double tk = house(M, n, 0, vb0, red, slot, &beta[1], &tau[0]);
        if (tk != 0.0) {
            fused_pass(M, n, -1, 0.0, nullptr, nullptr, tk, vb0, p);
            form_q(n - 1, tk, vb0, p, q, red, slot);
        }
        int cur = 0;
        for (int k = 0; k <= n - 3; ++k) {
            const int m = n - k - 1;
            z2* v = cur ? vb1 : vb0;
            z2* vn = cur ? vb0 : vb1;
            // column k+1, rows [k+1, n): rank-2 update of step k
            if (tk != 0.0) {
                const z2 q0 = q[0], v0 = v[0];
                z2* col = M + (k + 1) + static_cast<long>(k + 1) * n;
                for (int i = tid; i < m; i += nt) {
                    const z2 t1 = zmulc(v[i], q0), t2 = zmulc(q[i], v0);
                    col[i].x -= t1.x + t2.x;
                    col[i].y -= t1.y + t2.y;
                }
                __syncthreads();
            }
            if (tid == 0) d[k + 1] = M[(k + 1) + static_cast<long>(k + 1) * n].x;
            if (k == n - 3) {  // last 2 x 2 block
                if (tid == 0) {
                    z2* a = M + (n - 1) + static_cast<long>(n - 1) * n;
                    if (tk != 0.0) {
                        const z2 t1 = zmulc(v[1], q[1]), t2 = zmulc(q[1], v[1]);
                        a->x -= t1.x + t2.x;
                        a->y -= t1.y + t2.y;
                    }
                    d[n - 1] = a->x;
                    beta[n - 1] = M[(n - 1) + static_cast<long>(n - 2) * n];
                }
                break;
            }
            const double tn = house(M, n, k + 1, vn, red, slot, &beta[k + 2], &tau[k + 1]);
            if (tk != 0.0 || tn != 0.0) fused_pass(M, n, k, tk, v, q, tn, vn, p);
            if (tn != 0.0) form_q(m - 1, tn, vn, p, q, red, slot);
            tk = tn;
            cur ^= 1;
        }
    }
    __syncthreads();
    if (tid == 0) {
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
    if (LM)
        for (long e = tid; e < nn; e += nt) A[e] = M[e];
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
        q = (d[i] - x) - e2[i] / q;
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
__global__ __launch_bounds__(kThreads) void invit_kernel(int n, const double* __restrict__ d,
                                                         const double* __restrict__ e, double* w, int nvec,
                                                         double* __restrict__ Zt, double* __restrict__ U1,
                                                         double* __restrict__ U2, double* __restrict__ U3,
                                                         double* __restrict__ LO, int* __restrict__ SW) {
    __shared__ double red[32];
    __shared__ double sgl[4];
    int slot = 0;
    const int tid = threadIdx.x, nt = blockDim.x;
    if (tid < 64) {
        double gl, gu, tn, pivmin;
        tri_bounds(n, d, e, gl, gu, tn, pivmin);
        if (tid == 0) sgl[0] = tn;
    }
    if (tid == 0)  // eigenvalues from independent bisections: enforce ascending order
        for (int c = 1; c < nvec; ++c) w[c] = fmax(w[c], w[c - 1]);
    __syncthreads();
    const double tn = sgl[0];
    const double ctol = 1e-6 * tn;
    const double pivmin = fmax(kEps * tn, 1e-300);
    int maxrank = 0;
    for (int c = tid; c < nvec; c += nt) {
        int c0 = c;
        while (c0 > 0 && w[c0] - w[c0 - 1] <= ctol) --c0;
        maxrank = max(maxrank, c - c0);
    }
    const int rounds = static_cast<int>(block_max(static_cast<double>(maxrank), red, slot)) + 1;
    for (int r = 0; r < rounds; ++r) {
        for (int c = tid; c < nvec; c += nt) {
            int c0 = c;
            while (c0 > 0 && w[c0] - w[c0 - 1] <= ctol) --c0;
            if (c - c0 != r) continue;
            // shifts of a cluster kept apart by at least 10 eps |mu|
            double mu = w[c0];
            for (int cc = c0 + 1; cc <= c; ++cc) {
                double m = w[cc];
                if (m - mu < 10.0 * kEps * fabs(m)) m = mu + 10.0 * kEps * fmax(fabs(m), tn * kEps);
                mu = m;
            }
            auto at = [nvec, c](int i) { return static_cast<long>(i) * nvec + c; };
            // LU with partial pivoting of Tr - mu I: row i keeps (u1 diag, u2, u3 super-diagonals)
            for (int i = 0; i < n; ++i) {
                U1[at(i)] = d[i] - mu;
                U2[at(i)] = i + 1 < n ? e[i + 1] : 0.0;
                U3[at(i)] = 0.0;
            }
            for (int i = 0; i + 1 < n; ++i) {
                const double sub = e[i + 1];
                double u1 = U1[at(i)];
                if (fabs(u1) >= fabs(sub)) {
                    SW[at(i)] = 0;
                    if (fabs(u1) < pivmin) u1 = copysign(pivmin, u1 == 0.0 ? 1.0 : u1);
                    U1[at(i)] = u1;
                    const double f = sub / u1;
                    LO[at(i)] = f;
                    U1[at(i + 1)] -= f * U2[at(i)];
                } else {
                    SW[at(i)] = 1;
                    const double f = u1 / sub;
                    LO[at(i)] = f;
                    const double a2 = U2[at(i)], b2 = U1[at(i + 1)], b3 = U2[at(i + 1)];
                    U1[at(i)] = sub;
                    U2[at(i)] = b2;
                    U3[at(i)] = b3;
                    U1[at(i + 1)] = a2 - f * b2;
                    U2[at(i + 1)] = -f * b3;
                }
            }
            if (fabs(U1[at(n - 1)]) < pivmin) U1[at(n - 1)] = copysign(pivmin, U1[at(n - 1)] == 0.0 ? 1.0 : U1[at(n - 1)]);
            for (int i = 0; i < n; ++i) Zt[at(i)] = 1.0 + 0.1 * sin(1.3 * i + 0.7 * c);
            for (int it = 0; it < 3; ++it) {
                // forward: L^-1 with the recorded interchanges
                double xi = Zt[at(0)];
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
    for (int k = n - 3; k >= 0; --k) {
        const double t = tau[k];
        if (t == 0.0) continue;
        const z2* col = A + static_cast<long>(k) * n;
        z2 v[RPL];
        double sr = 0.0, si = 0.0;
#pragma unroll
        for (int r = 0; r < RPL; ++r) {
            const int i = lane + 64 * r;
            v[r] = i == k + 1 ? z2{1.0, 0.0} : (i > k + 1 && i < n ? col[i] : z2{0.0, 0.0});
            sr += v[r].x * y[r].x + v[r].y * y[r].y;
            si += v[r].x * y[r].y - v[r].y * y[r].x;
        }
        sr = t * wave_sum(sr);
        si = t * wave_sum(si);
#pragma unroll
        for (int r = 0; r < RPL; ++r) {
            y[r].x -= v[r].x * sr - v[r].y * si;
            y[r].y -= v[r].x * si + v[r].y * sr;
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
    size_t off[24];
    size_t total;
};
enum Buf { B_HF, B_SF, B_ZL, B_C, B_T1, B_Y, B_Z2, B_ZY, B_YB, B_G, B_Y2, B_D, B_E, B_TAU, B_W, B_ZT, B_U1, B_U2,
           B_U3, B_LO, B_SW, B_BETA, B_PH, B_ERR, B_COUNT };
Layout layout(int n, int nvec) {
    const size_t nn = static_cast<size_t>(n) * n * 16, nv = static_cast<size_t>(n) * nvec;
    const size_t sz[B_COUNT] = {nn, nn, nn, nn, nn, nn, nn, nn,                   // Hf Sf Zl C T1 Y Z2 ZY
                                nv * 16, static_cast<size_t>(nvec) * nvec * 16, nv * 16,  // Yb G Y2
                                size_t(n) * 8, size_t(n) * 8, size_t(n) * 8, size_t(nvec) * 8,  // d e tau w
                                nv * 8, nv * 8, nv * 8, nv * 8, nv * 8, nv * 4,  // Zt U1 U2 U3 LO SW
                                size_t(n) * 16, size_t(n) * 16, 64};              // beta phase err
    Layout L{};
    size_t o = 0;
    for (int i = 0; i < B_COUNT; ++i) {
        L.off[i] = o;
        o += (sz[i] + 255) & ~size_t(255);
    }
    L.total = o;
    return L;
}

size_t tridiag_lds(int n, bool lm) { return (4 * static_cast<size_t>(n) + 16 + (lm ? static_cast<size_t>(n) * n : 0)) * 16; }

}  // namespace

size_t subspace_eig_workspace(int n, int nvec) { return layout(std::max(n, 1), std::max(nvec, 1)).total; }

int subspace_eig(int n, const double* H, long ldh, const double* S, long lds, int nvec, double* w_host, double* X,
                 long ldx, void* work, size_t work_bytes, void* gemm_work, size_t gemm_bytes, size_t max_lds,
                 hipStream_t st, hipError_t* err, SubspaceEigStats* stats) {
    *err = hipSuccess;
    if (n <= 0 || nvec <= 0) return 0;
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

    hipLaunchKernelGGL(herm_prep, dim3(1), dim3(kThreads), 0, st, n, reinterpret_cast<const z2*>(H), ldh,
                       reinterpret_cast<const z2*>(S), lds, Hf, Sf, Zl, errd);
    if (!ok(hipGetLastError())) return -1;
    double sdev = 0.0;
    if (!ok(hipMemcpyAsync(&sdev, errd, 8, hipMemcpyDeviceToHost, st)) || !ok(hipStreamSynchronize(st))) return -1;
    if (stats) {
        stats->s_deviation = sdev;
        stats->newton_schulz_iterations = 0;
    }
    // first-order Loewdin factor error ~ (3/8) |E|_2^2 <= (3/8) (n max|E|)^2
    if (!(sdev * n <= 1e-7)) {
        if (!(sdev == sdev) || sdev > 1e6) return n + 1;
        // Newton-Schulz on S / c with c >= lambda_max(S): Y -> (S/c)^(1/2), Z -> (S/c)^(-1/2)
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
            if (!ok(hipGetLastError()) || !ok(hipMemcpyAsync(&r, errd, 8, hipMemcpyDeviceToHost, st)) ||
                !ok(hipStreamSynchronize(st)))
                return -1;
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
    }
    // C = Zl Hf Zl
    if (!gemm(0, n, n, n, Hf, n, Zl, n, T1, n) || !gemm(0, n, n, n, Zl, n, T1, n, C, n)) return -1;
    // tridiagonalisation
    double *d = D(B_D), *e = D(B_E), *tau = D(B_TAU), *w = D(B_W);
    z2 *beta = P(B_BETA), *ph = P(B_PH);
    const bool lm = tridiag_lds(n, true) <= max_lds;
    const size_t tl = tridiag_lds(n, lm);
    if (tl > max_lds) {
        *err = hipErrorInvalidValue;
        return -1;
    }
    static bool lds_raised = false;
    if (!lds_raised && tl > 65536) {
        if (!ok(hipFuncSetAttribute(reinterpret_cast<const void*>(&tridiag_kernel<true>),
                                    hipFuncAttributeMaxDynamicSharedMemorySize, static_cast<int>(max_lds))))
            return -1;
        lds_raised = true;
    }
    if (lm)
        hipLaunchKernelGGL(tridiag_kernel<true>, dim3(1), dim3(kThreads), tl, st, C, n, d, beta, tau, ph, e);
    else
        hipLaunchKernelGGL(tridiag_kernel<false>, dim3(1), dim3(kThreads), tl, st, C, n, d, beta, tau, ph, e);
    if (!ok(hipGetLastError())) return -1;
    hipLaunchKernelGGL(bisect_kernel, dim3(nvec), dim3(64), static_cast<size_t>(n) * 16, st, n, d, e, w);
    if (!ok(hipGetLastError())) return -1;
    hipLaunchKernelGGL(invit_kernel, dim3(1), dim3(kThreads), 0, st, n, d, e, w, nvec, D(B_ZT), D(B_U1), D(B_U2),
                       D(B_U3), D(B_LO), reinterpret_cast<int*>(base + L.off[B_SW]));
    if (!ok(hipGetLastError())) return -1;
    z2* Yb = P(B_YB);
    const int rpl = (n + 63) / 64;
#define Q3X_BT(R)                                                                                              \
    hipLaunchKernelGGL(backtransform_kernel<R>, dim3(nvec), dim3(64), 0, st, n, C, tau, ph, D(B_ZT), nvec, Yb)
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
    if (!ok(hipMemcpyAsync(w_host, w, static_cast<size_t>(nvec) * 8, hipMemcpyDeviceToHost, st)) ||
        !ok(hipStreamSynchronize(st)))
        return -1;
    for (int i = 0; i < nvec; ++i)
        if (!(w_host[i] == w_host[i])) return n + 1;
    return 0;
}

}  // namespace q3x
