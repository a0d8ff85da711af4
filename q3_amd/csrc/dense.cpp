// FP64 dense Hermitian eigensolver for the Davidson subspace problem (M <= a few hundred),
// on the host with OpenMP.  zhegv (itype 1) semantics as the reference's
// SerialSubspaceDiagonalizer (subspace/serial.hpp:26-136):
//   1. S = U^H U                                   (Cholesky, upper)
//   2. C = U^-H H U^-1                             (row / column triangular solves, AXPY form)
//   3. C = Q T Q^H                                  (Householder, real/imag split storage)
//   4. T = D Tr D^H, Tr = Z L Z^T                   (phase scaling, implicit-shift QL; each
//                                                    sweep's plane rotations are recorded and
//                                                    applied to row blocks of Z in parallel)
//   5. X = U^-1 Q D Z[:, :nvec]                      (only the wanted eigenvectors)
// All inner loops run over contiguous memory (column-major with AXPY updates) so they
// vectorise; the file is compiled with reassociation allowed (it is not bit-reproducible
// against LAPACK anyway, only accurate to rounding).
#include "dense.h"

#include <omp.h>

#include <algorithm>
#include <cmath>
#include <vector>

namespace q3x {

namespace {
inline size_t ix(int i, int j, int ld) { return static_cast<size_t>(i) + static_cast<size_t>(j) * ld; }

// Thread team of the Cholesky and triangular inverse of the Davidson orthogonalisation.  These
// run once per iteration on every solver thread at once (batch and server workers); full-size
// OpenMP teams from each of them oversubscribe the host, and their spinning threads starve the
// threads that drive the GPU (64 cfg4 k-points on 8 workers: 0.40 s with one thread per team,
// 0.74-0.80 s with full teams).  So: serial below n = 128, at most four threads above.
inline int small_team() { return std::max(1, std::min(4, omp_get_max_threads())); }

// Real and imaginary planes, column-major n x n, in one allocation.  The imaginary plane
// starts 2 KiB (mod 4 KiB) after the real one: with equal 4 KiB offsets, a load of im[i]
// right after a store to re[i] is taken by x86 cores as a possible store-to-load forward and
// stalls (4K aliasing), which slowed the AXPY loops below several times over.
struct Split {
    int n = 0;
    std::vector<double> buf;
    double* re = nullptr;
    double* im = nullptr;
    explicit Split(int n_) : n(n_) {
        const size_t nn = static_cast<size_t>(n_) * n_;
        const size_t shift = (256 + 512 - nn % 512) % 512;  // (nn + shift) * 8 = 2048 mod 4096
        buf.assign(2 * nn + shift + 8, 0.0);
        re = buf.data();
        im = buf.data() + nn + shift;
    }
    double* cr(int j) { return re + static_cast<size_t>(j) * n; }
    double* ci(int j) { return im + static_cast<size_t>(j) * n; }
};

// Upper Cholesky in place on split storage: A = U^H U.  Right-looking: after column j is
// final, the trailing rows j of columns i > j are updated with AXPY over k.
int chol_split(Split& A) {
    const int n = A.n;
    for (int j = 0; j < n; ++j) {
        const double* ar = A.cr(j);
        const double* ai = A.ci(j);
        double d = ar[j];
        for (int k = 0; k < j; ++k) d -= ar[k] * ar[k] + ai[k] * ai[k];
        if (!(d > 0.0)) return j + 1;
        d = std::sqrt(d);
        A.cr(j)[j] = d;
        A.ci(j)[j] = 0.0;
        const double inv = 1.0 / d;
#pragma omp parallel for schedule(static) num_threads(small_team()) if (n - j > 96)
        for (int i = j + 1; i < n; ++i) {
            const double* br = A.cr(i);
            const double* bi = A.ci(i);
            double sr = 0.0, si = 0.0;
            for (int k = 0; k < j; ++k) {  // conj(U[k][j]) U[k][i]
                sr += ar[k] * br[k] + ai[k] * bi[k];
                si += ar[k] * bi[k] - ai[k] * br[k];
            }
            A.cr(i)[j] = (A.cr(i)[j] - sr) * inv;
            A.ci(i)[j] = (A.ci(i)[j] - si) * inv;
        }
    }
    for (int j = 0; j < n; ++j)
        for (int i = j + 1; i < n; ++i) {
            A.cr(j)[i] = 0.0;
            A.ci(j)[i] = 0.0;
        }
    return 0;
}

// Implicit-shift QL on a real symmetric tridiagonal (d diagonal, e[i] couples i-1 and i on
// input, e[0] unused), eigenvalues in d, eigenvectors as columns of Z (column-major).
int ql_implicit(int n, std::vector<double>& d, std::vector<double>& e, std::vector<double>& Z) {
    for (int i = 1; i < n; ++i) e[i - 1] = e[i];
    e[n - 1] = 0.0;
    double f = 0.0, tst1 = 0.0;
    const double eps = 2.220446049250313e-16;
    std::vector<double> cs(n), sn(n);
    const int nt = std::max(1, std::min(omp_get_max_threads(), n / 32));
    for (int l = 0; l < n; ++l) {
        tst1 = std::max(tst1, std::abs(d[l]) + std::abs(e[l]));
        int m = l;
        while (m < n - 1 && std::abs(e[m]) > eps * tst1) ++m;
        if (m > l) {
            int iter = 0;
            do {
                if (++iter > 60) return l + 1;
                double g = d[l];
                double p = (d[l + 1] - g) / (2.0 * e[l]);
                double r = std::hypot(p, 1.0);
                if (p < 0) r = -r;
                d[l] = e[l] / (p + r);
                d[l + 1] = e[l] * (p + r);
                const double dl1 = d[l + 1];
                double h = g - d[l];
                for (int i = l + 2; i < n; ++i) d[i] -= h;
                f += h;
                p = d[m];
                double c = 1.0, c2 = 1.0, c3 = 1.0, s = 0.0, s2 = 0.0;
                const double el1 = e[l + 1];
                for (int i = m - 1; i >= l; --i) {
                    c3 = c2;
                    c2 = c;
                    s2 = s;
                    g = c * e[i];
                    h = c * p;
                    r = std::hypot(p, e[i]);
                    e[i + 1] = s * r;
                    s = e[i] / r;
                    c = p / r;
                    p = c * d[i] - s * g;
                    d[i + 1] = h + s * (c * g + s * d[i]);
                    cs[i] = c;
                    sn[i] = s;
                }
                // rotations i = m-1 .. l on columns (i, i+1) of Z, rows split over threads
                const int work = (m - l) * n;
#pragma omp parallel num_threads(nt) if (work > 8192)
                {
                    const int T = omp_get_num_threads(), t = omp_get_thread_num();
                    const int k0 = (n * t) / T, k1 = (n * (t + 1)) / T;
                    for (int i = m - 1; i >= l; --i) {
                        double* zi = Z.data() + static_cast<size_t>(i) * n;
                        double* zj = zi + n;
                        const double cc = cs[i], ss = sn[i];
#pragma omp simd
                        for (int k = k0; k < k1; ++k) {
                            const double a = zi[k], b = zj[k];
                            zj[k] = ss * a + cc * b;
                            zi[k] = cc * a - ss * b;
                        }
                    }
                }
                p = -s * s2 * c3 * el1 * e[l] / dl1;
                e[l] = s * p;
                d[l] = c * p;
            } while (std::abs(e[l]) > eps * tst1);
        }
        d[l] += f;
        e[l] = 0.0;
    }
    return 0;
}

// Eigenvalues only (the QL iteration above without accumulation), ascending.  O(n^2).
int ql_values(int n, std::vector<double> d, std::vector<double> e, std::vector<double>& out) {
    for (int i = 1; i < n; ++i) e[i - 1] = e[i];
    e[n - 1] = 0.0;
    double f = 0.0, tst1 = 0.0;
    const double eps = 2.220446049250313e-16;
    for (int l = 0; l < n; ++l) {
        tst1 = std::max(tst1, std::abs(d[l]) + std::abs(e[l]));
        int m = l;
        while (m < n - 1 && std::abs(e[m]) > eps * tst1) ++m;
        if (m > l) {
            int iter = 0;
            do {
                if (++iter > 60) return l + 1;
                double g = d[l];
                double p = (d[l + 1] - g) / (2.0 * e[l]);
                double r = std::hypot(p, 1.0);
                if (p < 0) r = -r;
                d[l] = e[l] / (p + r);
                d[l + 1] = e[l] * (p + r);
                const double dl1 = d[l + 1];
                double h = g - d[l];
                for (int i = l + 2; i < n; ++i) d[i] -= h;
                f += h;
                p = d[m];
                double c = 1.0, c2 = 1.0, c3 = 1.0, s = 0.0, s2 = 0.0;
                const double el1 = e[l + 1];
                for (int i = m - 1; i >= l; --i) {
                    c3 = c2;
                    c2 = c;
                    s2 = s;
                    g = c * e[i];
                    h = c * p;
                    r = std::hypot(p, e[i]);
                    e[i + 1] = s * r;
                    s = e[i] / r;
                    c = p / r;
                    p = c * d[i] - s * g;
                    d[i + 1] = h + s * (c * g + s * d[i]);
                }
                p = -s * s2 * c3 * el1 * e[l] / dl1;
                e[l] = s * p;
                d[l] = c * p;
            } while (std::abs(e[l]) > eps * tst1);
        }
        d[l] += f;
        e[l] = 0.0;
    }
    out = d;
    std::sort(out.begin(), out.end());
    return 0;
}

// Eigenvectors of the real symmetric tridiagonal (d, e[i] couples i-1 and i) for eigenvalues
// lam[0..nvec), by inverse iteration on (T - lam I) factored with partial pivoting, with
// modified Gram-Schmidt against earlier vectors of the same cluster (eigenvalues closer than
// 1e-3 |T|_1, the criterion of LAPACK's dstein).  Z: n x nvec column-major.
void tridiag_vectors(int n, const std::vector<double>& d, const std::vector<double>& e, const double* lam, int nvec,
                     std::vector<double>& Z) {
    double tnorm = 0.0;
    for (int i = 0; i < n; ++i)
        tnorm = std::max(tnorm, std::abs(d[i]) + (i > 0 ? std::abs(e[i]) : 0.0) + (i + 1 < n ? std::abs(e[i + 1]) : 0.0));
    const double ortol = 1e-3 * tnorm;
    const double eps = 2.220446049250313e-16;
    const double pivmin = std::max(eps * tnorm, 1e-300);
    // clusters of consecutive eigenvalues
    std::vector<int> cstart;
    for (int c = 0; c < nvec; ++c)
        if (c == 0 || lam[c] - lam[c - 1] > ortol) cstart.push_back(c);
    cstart.push_back(nvec);
    Z.assign(static_cast<size_t>(n) * nvec, 0.0);
    const int nclusters = static_cast<int>(cstart.size()) - 1;
#pragma omp parallel
    {
        std::vector<double> u1(n), u2(n), u3(n), lo(n), x(n);
        std::vector<int> swp(n);
#pragma omp for schedule(dynamic, 1)
        for (int cl = 0; cl < nclusters; ++cl) {
            const int c0 = cstart[cl], c1 = cstart[cl + 1];
            double prev = 0.0;
            for (int c = c0; c < c1; ++c) {
                double mu = lam[c];
                // separate exactly equal shifts inside a cluster slightly (dstein does the same)
                if (c > c0 && mu - prev < 10.0 * eps * std::abs(mu)) mu = prev + 10.0 * eps * std::max(std::abs(mu), tnorm * eps);
                prev = mu;
                // LU with partial pivoting of T - mu I: rows kept as (u1 diag, u2 super, u3 super2)
                for (int i = 0; i < n; ++i) {
                    u1[i] = d[i] - mu;
                    u2[i] = (i + 1 < n) ? e[i + 1] : 0.0;
                    u3[i] = 0.0;
                }
                for (int i = 0; i + 1 < n; ++i) {
                    const double sub = e[i + 1];  // T[i+1][i]
                    if (std::abs(u1[i]) >= std::abs(sub)) {
                        swp[i] = 0;
                        if (std::abs(u1[i]) < pivmin) u1[i] = pivmin;
                        const double f = sub / u1[i];
                        lo[i] = f;
                        u1[i + 1] -= f * u2[i];
                        // u3[i] stays 0; row i+1 super stays u2[i+1]
                    } else {
                        swp[i] = 1;
                        const double f = u1[i] / sub;
                        lo[i] = f;
                        // swap rows i and i+1
                        const double a1 = u1[i], a2 = u2[i];
                        const double b2 = u1[i + 1], b3 = u2[i + 1];
                        u1[i] = sub;
                        u2[i] = b2;
                        u3[i] = b3;
                        u1[i + 1] = a2 - f * b2;
                        u2[i + 1] = -f * b3;
                        (void)a1;
                    }
                }
                if (std::abs(u1[n - 1]) < pivmin) u1[n - 1] = pivmin;
                // deterministic start vector
                for (int i = 0; i < n; ++i) x[i] = 1.0 + 0.1 * std::sin(1.3 * i + 0.7 * c);
                for (int itn = 0; itn < 3; ++itn) {
                    // forward: apply L^-1 with the recorded swaps
                    for (int i = 0; i + 1 < n; ++i) {
                        if (swp[i]) std::swap(x[i], x[i + 1]);
                        x[i + 1] -= lo[i] * x[i];
                    }
                    // backward: U x = y
                    for (int i = n - 1; i >= 0; --i) {
                        double s = x[i];
                        if (i + 1 < n) s -= u2[i] * x[i + 1];
                        if (i + 2 < n) s -= u3[i] * x[i + 2];
                        x[i] = s / u1[i];
                    }
                    // orthogonalise against earlier vectors of the cluster
                    for (int q = c0; q < c; ++q) {
                        const double* zq = Z.data() + static_cast<size_t>(q) * n;
                        double dot = 0.0;
                        for (int i = 0; i < n; ++i) dot += zq[i] * x[i];
                        for (int i = 0; i < n; ++i) x[i] -= dot * zq[i];
                    }
                    double nrm = 0.0;
                    for (int i = 0; i < n; ++i) nrm += x[i] * x[i];
                    nrm = std::sqrt(nrm);
                    const double inv = nrm > 0.0 ? 1.0 / nrm : 0.0;
                    for (int i = 0; i < n; ++i) x[i] *= inv;
                }
                std::copy(x.begin(), x.end(), Z.begin() + static_cast<size_t>(c) * n);
            }
        }
    }
}

}  // namespace

int chol_upper(int n, zc* A, int lda) {
    Split S(n);
    for (int j = 0; j < n; ++j)
        for (int i = 0; i <= j; ++i) {
            S.cr(j)[i] = A[ix(i, j, lda)].real();
            S.ci(j)[i] = A[ix(i, j, lda)].imag();
        }
    const int info = chol_split(S);
    if (info) return info;
    for (int j = 0; j < n; ++j)
        for (int i = 0; i < n; ++i) A[ix(i, j, lda)] = (i <= j) ? zc(S.cr(j)[i], S.ci(j)[i]) : zc(0, 0);
    return 0;
}

void tri_inv_upper(int n, zc* U, int ldu) {
    // columns of U^-1 are independent: U x_j = e_j, backward AXPY form over columns of U
    std::vector<double> Ur(static_cast<size_t>(n) * n), Ui(static_cast<size_t>(n) * n);
    for (int j = 0; j < n; ++j)
        for (int i = 0; i <= j; ++i) {
            Ur[ix(i, j, n)] = U[ix(i, j, ldu)].real();
            Ui[ix(i, j, n)] = U[ix(i, j, ldu)].imag();
        }
    std::vector<double> Xr(static_cast<size_t>(n) * n, 0.0), Xi(static_cast<size_t>(n) * n, 0.0);
#pragma omp parallel for schedule(dynamic, 4) num_threads(small_team()) if (n >= 128)
    for (int j = 0; j < n; ++j) {
        double* xr = Xr.data() + static_cast<size_t>(j) * n;
        double* xi = Xi.data() + static_cast<size_t>(j) * n;
        xr[j] = 1.0;
        for (int k = j; k >= 0; --k) {
            const double* ur = Ur.data() + static_cast<size_t>(k) * n;
            const double* ui = Ui.data() + static_cast<size_t>(k) * n;
            const double dr = ur[k], di = ui[k], dd = dr * dr + di * di;
            const double vr = (xr[k] * dr + xi[k] * di) / dd, vi = (xi[k] * dr - xr[k] * di) / dd;
            xr[k] = vr;
            xi[k] = vi;
#pragma omp simd
            for (int i = 0; i < k; ++i) {
                xr[i] -= ur[i] * vr - ui[i] * vi;
                xi[i] -= ur[i] * vi + ui[i] * vr;
            }
        }
    }
    for (int j = 0; j < n; ++j)
        for (int i = 0; i < n; ++i)
            U[ix(i, j, ldu)] = (i <= j) ? zc(Xr[ix(i, j, n)], Xi[ix(i, j, n)]) : zc(0, 0);
}

int hegv(int n, const zc* H, int ldh, const zc* S, int lds, double* w, zc* X, int ldx, int nvec) {
    if (n <= 0) return 0;
    nvec = std::min(nvec, n);
    // 1. Cholesky of S (upper triangle)
    Split U(n);
    for (int j = 0; j < n; ++j)
        for (int i = 0; i <= j; ++i) {
            U.cr(j)[i] = S[ix(i, j, lds)].real();
            U.ci(j)[i] = S[ix(i, j, lds)].imag();
        }
    int info = chol_split(U);
    if (info) return n + info;
    // Ut: U transposed (row k of U contiguous), used by both triangular solves
    Split Ut(n);
    for (int j = 0; j < n; ++j)
        for (int k = 0; k <= j; ++k) {
            Ut.cr(k)[j] = U.cr(j)[k];
            Ut.ci(k)[j] = U.ci(j)[k];
        }
    // 2a. W = Hfull U^-1, computed as W^T rows... equivalently W^H = U^-H Hfull: solve column j
    //     of W^H: U^H y = Hfull[:, j]^* ... we solve (W row i) U = (H row i) via columns of H^T.
    //     With H Hermitian, row i of H = conj(column i of H), so W row i = conj(y_i) where
    //     U^H y_i = column i of H.  Forward AXPY over k with row k of U (= Ut column k).
    Split C(n);
    Split Wt(n);  // column i of Wt = conj(row i of W) = y_i
#pragma omp parallel for schedule(static)
    for (int i = 0; i < n; ++i) {
        double* yr = Wt.cr(i);
        double* yi = Wt.ci(i);
        for (int r = 0; r < n; ++r) {
            const zc hv = (r <= i) ? H[ix(r, i, ldh)] : std::conj(H[ix(i, r, ldh)]);
            yr[r] = hv.real();
            yi[r] = (r == i) ? 0.0 : hv.imag();
        }
        for (int k = 0; k < n; ++k) {
            const double* ur = Ut.cr(k);  // ur[q] = U[k][q]
            const double* ui = Ut.ci(k);
            const double dkk = ur[k];
            const double vr = yr[k] / dkk, vi = yi[k] / dkk;
            yr[k] = vr;
            yi[k] = vi;
            // y[q] -= conj(U[k][q]) v  for q > k
#pragma omp simd
            for (int q = k + 1; q < n; ++q) {
                yr[q] -= ur[q] * vr + ui[q] * vi;
                yi[q] -= ur[q] * vi - ui[q] * vr;
            }
        }
    }
    // 2b. C = U^-H W: column j of C solves U^H c = W[:, j]; W[r][j] = conj(Wt[j][r])
#pragma omp parallel for schedule(static)
    for (int j = 0; j < n; ++j) {
        double* cr = C.cr(j);
        double* ci = C.ci(j);
        for (int r = 0; r < n; ++r) {
            cr[r] = Wt.cr(r)[j];
            ci[r] = -Wt.ci(r)[j];
        }
        for (int k = 0; k < n; ++k) {
            const double* ur = Ut.cr(k);
            const double* ui = Ut.ci(k);
            const double dkk = ur[k];
            const double vr = cr[k] / dkk, vi = ci[k] / dkk;
            cr[k] = vr;
            ci[k] = vi;
#pragma omp simd
            for (int q = k + 1; q < n; ++q) {
                cr[q] -= ur[q] * vr + ui[q] * vi;
                ci[q] -= ur[q] * vi - ui[q] * vr;
            }
        }
    }
    // Hermitian symmetrisation against rounding
    for (int j = 0; j < n; ++j) {
        C.ci(j)[j] = 0.0;
        for (int i = 0; i < j; ++i) {
            const double ar = 0.5 * (C.cr(j)[i] + C.cr(i)[j]);
            const double ai = 0.5 * (C.ci(j)[i] - C.ci(i)[j]);
            C.cr(j)[i] = ar;
            C.ci(j)[i] = ai;
            C.cr(i)[j] = ar;
            C.ci(i)[j] = -ai;
        }
    }
    // 3. Householder tridiagonalisation.  Reflector k acts on rows k+1..n-1; its vector
    // (v0 = 1) is kept in column k below the sub-diagonal, tau[k] real.
    // One parallel region for all n-2 steps.  Column j of the trailing block belongs to
    // thread j mod T for the whole reduction, and both halves of a step touch only owned
    // columns: p_j = t * (column j)^H v (the block is Hermitian, so this is row j of B v), then
    // column j -= v conj(q_j) + q conj(v_j).  The block therefore stays in the owners' caches;
    // a step costs two barriers.
    std::vector<double> tau(n, 0.0), vr(n), vi(n), pr(n), pi(n);
    std::vector<double> offr(n, 0.0), offi(n, 0.0);
    const int nthreads = std::max(1, std::min(omp_get_max_threads(), n / 16));
#pragma omp parallel num_threads(nthreads)
    {
        const int T = omp_get_num_threads(), tid = omp_get_thread_num();
        for (int k = 0; k < n - 2; ++k) {
            const int len = n - k - 1, off = k + 1;
            double* xr = C.cr(k) + off;
            double* xi = C.ci(k) + off;
#pragma omp single
            {
                double xn2 = 0.0;
                for (int i = 0; i < len; ++i) xn2 += xr[i] * xr[i] + xi[i] * xi[i];
                const double xnorm = std::sqrt(xn2);
                if (xnorm == 0.0) {
                    tau[k] = 0.0;
                } else {
                    const double aabs = std::hypot(xr[0], xi[0]);
                    const double phr = aabs == 0.0 ? 1.0 : xr[0] / aabs, phi = aabs == 0.0 ? 0.0 : xi[0] / aabs;
                    const double br = -phr * xnorm, bi = -phi * xnorm;  // beta = -phase * |x|
                    const double v0r = xr[0] - br, v0i = xi[0] - bi;
                    const double v0n = v0r * v0r + v0i * v0i;
                    vr[0] = 1.0;
                    vi[0] = 0.0;
                    double vn2 = 1.0;
                    for (int i = 1; i < len; ++i) {
                        vr[i] = (xr[i] * v0r + xi[i] * v0i) / v0n;
                        vi[i] = (xi[i] * v0r - xr[i] * v0i) / v0n;
                        vn2 += vr[i] * vr[i] + vi[i] * vi[i];
                    }
                    for (int i = 1; i < len; ++i) {  // reflector kept in column k
                        xr[i] = vr[i];
                        xi[i] = vi[i];
                    }
                    tau[k] = 2.0 / vn2;
                    offr[off] = br;
                    offi[off] = bi;
                }
            }  // implicit barrier: v, tau visible
            const double t = tau[k];
            if (t != 0.0) {
                // p = t * B v over owned columns
                for (int j = (tid - off % T + T) % T; j < len; j += T) {
                    const double* br_ = C.cr(off + j) + off;
                    const double* bi_ = C.ci(off + j) + off;
                    double sr = 0.0, si = 0.0;
#pragma omp simd reduction(+ : sr, si)
                    for (int i = 0; i < len; ++i) {  // conj(B[i][j]) v[i]
                        sr += br_[i] * vr[i] + bi_[i] * vi[i];
                        si += br_[i] * vi[i] - bi_[i] * vr[i];
                    }
                    pr[j] = t * sr;
                    pi[j] = t * si;
                }
#pragma omp barrier
                double vhp = 0.0;  // v^H p is real; every thread forms it in the same order
                for (int i = 0; i < len; ++i) vhp += vr[i] * pr[i] + vi[i] * pi[i];
                const double K = 0.5 * t * vhp;
                // B -= v q^H + q v^H with q = p - K v, owned columns
                for (int j = (tid - off % T + T) % T; j < len; j += T) {
                    double* cr = C.cr(off + j) + off;
                    double* ci = C.ci(off + j) + off;
                    const double wr = vr[j], wi = vi[j];
                    const double qr = pr[j] - K * wr, qi = pi[j] - K * wi;
#pragma omp simd
                    for (int i = 0; i < len; ++i) {
                        const double sr = pr[i] - K * vr[i], si = pi[i] - K * vi[i];
                        cr[i] -= (vr[i] * qr + vi[i] * qi) + (sr * wr + si * wi);
                        ci[i] -= (vi[i] * qr - vr[i] * qi) + (si * wr - sr * wi);
                    }
                }
            }
#pragma omp barrier
        }
    }
    for (int k = 0; k < n - 2; ++k)
        if (tau[k] == 0.0) {  // x was zero: sub-diagonal entry is the (zero) matrix element
            offr[k + 1] = C.cr(k)[k + 1];
            offi[k + 1] = C.ci(k)[k + 1];
        }
    if (n >= 2) {
        offr[n - 1] = C.cr(n - 2)[n - 1];
        offi[n - 1] = C.ci(n - 2)[n - 1];
    }
    // 4. real tridiagonal via phases, QL
    std::vector<double> d(n), e(n, 0.0), phr(n, 1.0), phi(n, 0.0);
    for (int i = 0; i < n; ++i) d[i] = C.cr(i)[i];
    for (int i = 1; i < n; ++i) {
        const double r = std::hypot(offr[i], offi[i]);
        e[i] = r;
        if (r == 0.0) {
            phr[i] = phr[i - 1];
            phi[i] = phi[i - 1];
        } else {
            const double ur = offr[i] / r, ui = offi[i] / r;
            phr[i] = phr[i - 1] * ur - phi[i - 1] * ui;
            phi[i] = phr[i - 1] * ui + phi[i - 1] * ur;
        }
    }
    // eigenvalues by QL without accumulation, eigenvectors of the wanted ones by inverse
    // iteration (QL with full accumulation costs ~4.5 n^3 and is only kept as ql_implicit for
    // small problems where all vectors are wanted)
    std::vector<double> Z;
    std::vector<int> order(n);
    for (int i = 0; i < n; ++i) order[i] = i;
    if (nvec == n && n <= 64) {
        Z.assign(static_cast<size_t>(n) * n, 0.0);
        for (int i = 0; i < n; ++i) Z[ix(i, i, n)] = 1.0;
        std::vector<double> dd = d, ee = e;
        info = ql_implicit(n, dd, ee, Z);
        if (info) return info;
        std::stable_sort(order.begin(), order.end(), [&dd](int a, int b) { return dd[a] < dd[b]; });
        for (int i = 0; i < n; ++i) w[i] = dd[order[i]];
    } else {
        std::vector<double> vals;
        info = ql_values(n, d, e, vals);
        if (info) return info;
        for (int i = 0; i < n; ++i) w[i] = vals[i];
        tridiag_vectors(n, d, e, w, nvec, Z);
    }
    // 5. Y = Q D Z[:, wanted], then X = U^-1 Y (backward AXPY over columns of U)
#pragma omp parallel
    {
        std::vector<double> yr(n), yi(n);
#pragma omp for schedule(static)
        for (int c = 0; c < nvec; ++c) {
            const double* z = Z.data() + static_cast<size_t>(nvec == n && n <= 64 ? order[c] : c) * n;
            for (int i = 0; i < n; ++i) {
                yr[i] = phr[i] * z[i];
                yi[i] = phi[i] * z[i];
            }
            for (int k = n - 3; k >= 0; --k) {
                if (tau[k] == 0.0) continue;
                const int len = n - k - 1;
                const double* hr = C.cr(k) + k + 1;
                const double* hi = C.ci(k) + k + 1;
                double* ar = yr.data() + k + 1;
                double* ai = yi.data() + k + 1;
                double sr = ar[0], si = ai[0];
#pragma omp simd reduction(+ : sr, si)
                for (int i = 1; i < len; ++i) {
                    sr += hr[i] * ar[i] + hi[i] * ai[i];
                    si += hr[i] * ai[i] - hi[i] * ar[i];
                }
                sr *= tau[k];
                si *= tau[k];
                ar[0] -= sr;
                ai[0] -= si;
#pragma omp simd
                for (int i = 1; i < len; ++i) {
                    ar[i] -= hr[i] * sr - hi[i] * si;
                    ai[i] -= hr[i] * si + hi[i] * sr;
                }
            }
            for (int k = n - 1; k >= 0; --k) {
                const double* ur = U.cr(k);
                const double* ui = U.ci(k);
                const double dkk = ur[k];
                const double vr_ = yr[k] / dkk, vi_ = yi[k] / dkk;
                yr[k] = vr_;
                yi[k] = vi_;
#pragma omp simd
                for (int i = 0; i < k; ++i) {
                    yr[i] -= ur[i] * vr_ - ui[i] * vi_;
                    yi[i] -= ur[i] * vi_ + ui[i] * vr_;
                }
            }
            zc* x = X + static_cast<size_t>(c) * ldx;
            for (int i = 0; i < n; ++i) x[i] = zc(yr[i], yi[i]);
        }
    }
    return 0;
}

}  // namespace q3x

// C ABI: the subspace eigensolver on host arrays (stands in for
// GPUEigensolver::generalized_hermitian_eigenvalue_solve, math/linear_algebra/gpu/lapack/init.hpp).
extern "C" int q3x_hegv(int n, const double* H, const double* S, double* w, double* X, int nvec) {
    if (n < 0 || nvec < 0 || (n > 0 && (!H || !S || !w || (nvec > 0 && !X)))) return 1;
    return q3x::hegv(n, reinterpret_cast<const q3x::zc*>(H), n, reinterpret_cast<const q3x::zc*>(S), n, w,
                     reinterpret_cast<q3x::zc*>(X), n, nvec) == 0
               ? 0
               : 4;
}
