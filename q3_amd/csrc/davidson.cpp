// Block Davidson with a device-resident basis (PlanewaveDavidsonDiagonalizer::solve,
// davidson.hpp:324-841, with the subspace manager of subspace_manager/serial.hpp).
//
// The basis V, H V, the correction block, the subspace matrices H_sub / S_sub and their
// eigenvectors live in HBM; every product over the plane-wave dimension is an MFMA complex
// GEMM and the M x M generalized eigenproblem is solved on the device (subspace_eig.hip).
// The host keeps what the reference also decides on the host: the eigenvalue-change
// convergence test, the choice of corrections, the n_c x n_c Cholesky of the Cholesky-QR and
// the collapse decision.  Control flow and constants are the reference's:
//   M_max = max(min(factor*nb, npw), nb)                           davidson.hpp:399-404
//   iteration 0 never converges; occupied |de| < tol, empty < max(5 tol, 1e-5)  :607-629
//   corrections r = (HV - e V) c for unconverged bands, Teter preconditioner  serial.hpp:995-1156
//   projection + Cholesky-QR, modified Gram-Schmidt fallback (norm^2 > 1e-14)  serial.hpp:41-553
//   collapse to nb Ritz vectors when the basis is full, rebuild H_sub/S_sub   serial.hpp:825-945
//   zhegv failure -> forced collapse and retry, else previous eigenvalues   davidson.hpp:520-590
//
// Mixed precision: the H-apply runs in complex64 (FFT, V multiply, projector GEMMs), and so do
// the products that form basis vectors (rotations, corrections) when the tolerance is at least
// 1e-6 Ry; the basis storage, Gram products, Cholesky and Rayleigh-Ritz stay fp64.  With
// final_fp64_sweep the
// converged mixed basis is re-applied with the fp64 H and the iteration continues in fp64
// until the fp64 test passes, so the returned eigenpairs satisfy the fp64 criterion.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <stdexcept>
#include <vector>

#include "aux_kernels.h"
#include "dense.h"
#include "gamma_kernels.h"
#include "runtime.h"
#include "subspace_eig.h"
#include "zgemm.h"

namespace q3x {
namespace {

constexpr double kNormThreshold = 1e-14;  // iterative_solver_norm_threshold (tolerances.hpp:39)
constexpr double kEmptyFactor = 5.0;      // tolerances.hpp:65
constexpr double kEmptyFloor = 1e-5;      // tolerances.hpp:66
// mixed phase with complex64 basis products: the basis is orthonormal to their rounding, so
// S = V^H V deviates from I by about 1e-6 (n max|S - I| up to a few 1e-4 at M = 320); the
// first-order S^(-1/2) (error (3/8) (n max|S - I|)^2 <= 1e-7) is used up to 5e-4 instead of a
// Newton-Schulz iteration and a second solve
constexpr double kLowdinBoundMixed = 5e-4;
// complex64 basis products only for tolerances above their rounding floor: below it the mixed
// phase cannot meet the |de| test and runs on until its cap (cfg1 SCF, whose Davidson tolerance
// falls to 1e-13: 3,949 -> 4,423 iterations with them at every tolerance)
constexpr double kMixedProductsMinTol = 1e-6;

using clk = std::chrono::steady_clock;

struct Solver {
    explicit Solver(q3x_hamiltonian h_)
        : h(h_), st(h_->ctx->stream), ws(h_->ctx->dav), V(ws.V), HV(ws.HV), R(ws.R), C(ws.C), G(ws.G),
          Hs_d(ws.Hs), Ss_d(ws.Ss), X_d(ws.X), Xc_d(ws.Xc), eps_d(ws.eps), negeps_d(ws.negeps), work(ws.work),
          norms_d(ws.norms), idx_d(ws.idx), eigw(ws.eigw) {}
    q3x_hamiltonian h;
    hipStream_t st;
    int npw = 0, nb = 0, nocc = 0, mmax = 0;
    int precision = Q3X_FP64;
    // Gamma trick: half-sphere columns of real functions; every product over plane waves is the
    // full-sphere real one and G = 0 stays real (the gamma paths of subspace_manager/serial.hpp)
    bool gamma = false;
    // buffers of the context's workspace.  X_d: subspace eigenvectors (n x nvec, ld n) of the
    // last diag(); Xc_d: their columns for the corrections; idx_d / eps_d / negeps_d: band
    // indices and energies of the corrections
    DavidsonWorkspace& ws;
    DevBuf &V, &HV, &R, &C, &G, &Hs_d, &Ss_d, &X_d, &Xc_d, &eps_d, &negeps_d, &work, &norms_d, &idx_d, &eigw;
    size_t work_bytes = size_t(256) << 20;
    int h_applications = 0;
    SubspaceEigStats eig_stats;  // barrier_redos accumulates over the solve's eigensolves
    int mgs_fallbacks = 0, recovery_collapses = 0, previous_eigenvalue_keeps = 0;
    double stage[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    // With profiling on, every stage ends with a stream synchronisation so that stage_ms are
    // device times; otherwise the solve synchronises only where the host needs device data
    // (the Gram block of the Cholesky-QR, the subspace eigenvalues) and stage_ms measure the
    // host side.
    bool timing = false;
    bool stage_sync() { return !timing || sync(); }
    q3x_status status = Q3X_OK;

    // Current basis and its H image, and the other buffers of the ping-pong (rotate() swaps them)
    double *vbase = nullptr, *hvbase = nullptr, *valt = nullptr, *hvalt = nullptr;
    double* Vc(int c) { return vbase + 2L * c * npw; }
    double* HVc(int c) { return hvbase + 2L * c * npw; }

    bool ok(hipError_t e, const char* what) {
        if (hip_ok(e, what)) return true;
        status = Q3X_ERR_DEVICE;
        return false;
    }
    bool sync() { return ok(hipStreamSynchronize(st), "hipStreamSynchronize"); }
    HostStage& pin() { return h->ctx->pin; }
    double now_ms() { return std::chrono::duration<double, std::milli>(clk::now().time_since_epoch()).count(); }

    // mixed phase: the products that form basis vectors (A B: projections out of the corrections,
    // rotations, collapse) run on complex64 MFMA from the complex128 arrays; the Gram products
    // (A^H B: S, H_sub, projections onto the basis) stay FP64, so Rayleigh-Ritz stays exact on
    // whatever subspace the basis spans.  cfg2: 0.198 -> 0.194 s per solve, same errors; the
    // Gram products in complex64 as well: 0.235 s (the complex64 A^H B tiles are no faster here)
    bool fp32_products = false;
    // C (m x n) = alpha * op(A) * B + beta * C over complex128 arrays
    bool gemm(int opa, int m, int n, int k, double alpha, const void* A, long lda, const void* B, long ldb,
              double beta, void* Cm, long ldc) {
        const GemmTypes t = fp32_products && opa == 0 ? GemmTypes::D_D_F_D : GemmTypes::D_D_D_D;
        return ok(zgemm(t, opa, m, n, k, alpha, A, lda, B, ldb, beta, Cm, ldc, work.get(), work_bytes, st),
                  "Davidson GEMM");
    }

    // C (m x n) = A^H B over the plane waves; at Gamma 2 Re(A^H B) less the G = 0 term, real
    // (subspace_manager/serial.hpp:73-110, 228-262, 575-640)
    bool gram(int m, int n, const double* A, const double* B, void* Cm, long ldc) {
        if (!gemm(1, m, n, npw, 1.0, A, npw, B, npw, 0.0, Cm, ldc)) return false;
        return !gamma || ok(launch_gamma_gram_fix(m, n, A, npw, B, npw, static_cast<double*>(Cm), ldc, st),
                            "Gamma Gram");
    }
    // enforce_gamma_g0_real (subspace_manager/serial.hpp:1189-1200) on ncol columns
    bool g0_real(double* X, int ncol) { return !gamma || ok(launch_g0_real(ncol, X, npw, st), "Gamma G = 0"); }

    bool apply(int first, int n, int prec) {
        if (n <= 0) return true;
        const double t0 = now_ms();
        q3x_status s = hamiltonian_apply_device(h, Vc(first), npw, HVc(first), npw, n, prec);
        if (s != Q3X_OK) {
            status = s;
            return false;
        }
        h_applications += n;
        if (!stage_sync()) return false;
        stage[2] += now_ms() - t0;
        return true;
    }

    // H_sub[:total, start:start+nnew] = V^H HV_new ; S_sub likewise (serial.hpp:555-678).
    // Upper triangles of the leading total x total blocks are what diag() reads.
    bool update(int start, int nnew, int total) {
        if (nnew <= 0) return true;
        const double t0 = now_ms();
        double* hs = Hs_d.as<double>() + 2L * start * mmax;
        double* ss = Ss_d.as<double>() + 2L * start * mmax;
        if (!gram(total, nnew, Vc(0), HVc(start), hs, mmax)) return false;
        if (!gram(total, nnew, Vc(0), Vc(start), ss, mmax)) return false;
        if (!stage_sync()) return false;
        stage[3] += now_ms() - t0;
        return true;
    }

    // Lowest nvec eigenpairs of the leading n x n pencil; eigenvalues into w (host), vectors
    // into X_d (n x nvec, ld n).  Returns 0, n + 1 (S not positive definite) or -1.
    int diag(int n, std::vector<double>& w, int nvec) {
        const double t0 = now_ms();
        nvec = std::max(1, std::min(nvec, n));
        w.assign(std::max(n, mmax), 0.0);
        hipError_t e = hipSuccess;
        const int info = subspace_eig(n, Hs_d.as<double>(), mmax, Ss_d.as<double>(), mmax, nvec, w.data(),
                                      X_d.as<double>(), n, eigw.get(), eigw.size(), work.get(), work_bytes,
                                      h->ctx->max_lds, h->ctx->num_cus, st, &e, &eig_stats, &h->ctx->prof, &pin(),
                                      fp32_products ? kLowdinBoundMixed : 1e-7);
        if (info < 0) {
            ok(e == hipSuccess ? hipErrorUnknown : e, "subspace eigensolver");
            return -1;
        }
        // Gamma: the pencil is real symmetric and its eigenvectors real (dsygv, subspace/serial.hpp:33-90)
        if (gamma && info == 0 && !ok(launch_zero_imag(n, nvec, X_d.as<double>(), n, st), "Gamma real eigenvectors"))
            return -1;
        stage[4] += now_ms() - t0;
        return info;
    }

    // Orthogonalise candidates against V[:, :start] and among themselves (serial.hpp:41-553); the
    // result lands in V[:, start:start+nc].  The candidates are V[:, start:start+nc] itself, or,
    // in_r, the workspace R (the corrections are formed there, so the final product cand U^-1 is
    // written straight into V instead of into R and copied back).  Returns the accepted count
    // (compacted), -1 on device error.
    int orthogonalize(int start, int nc, bool in_r = false) {
        if (nc <= 0) return 0;
        const double t0 = now_ms();
        double* cand = in_r ? R.as<double>() : Vc(start);
        if (start > 0) {
            if (!gram(start, nc, Vc(0), cand, C.get(), start)) return -1;
            if (!gemm(0, npw, nc, start, -1.0, Vc(0), npw, C.get(), start, 1.0, cand, npw)) return -1;
        }
        if (!gram(nc, nc, cand, cand, G.get(), nc)) return -1;
        std::vector<zc> Gh(static_cast<size_t>(nc) * nc);
        if (!ok(pin().d2h(Gh.data(), G.get(), Gh.size() * 16, st), "D2H Gram")) return -1;
        int accepted = nc;
        if (chol_upper(nc, Gh.data(), nc) == 0) {
            tri_inv_upper(nc, Gh.data(), nc);  // cand <- cand U^-1 (the ztrsm R,U,N,N of serial.hpp:337-348)
            if (!ok(pin().h2d(G.get(), Gh.data(), Gh.size() * 16, st), "H2D U^-1")) return -1;
            if (in_r) {
                if (!gemm(0, npw, nc, nc, 1.0, cand, npw, G.get(), nc, 0.0, Vc(start), npw)) return -1;
            } else {
                if (!gemm(0, npw, nc, nc, 1.0, cand, npw, G.get(), nc, 0.0, R.get(), npw)) return -1;
                if (!ok(hipMemcpyAsync(cand, R.get(), static_cast<size_t>(npw) * nc * 16, hipMemcpyDeviceToDevice,
                                       st),
                        "D2D orthonormal block"))
                    return -1;
            }
        } else {
            // modified Gram-Schmidt fallback (serial.hpp:370-552), on the candidates moved into V
            if (in_r) {
                if (!ok(hipMemcpyAsync(Vc(start), cand, static_cast<size_t>(npw) * nc * 16, hipMemcpyDeviceToDevice,
                                       st),
                        "D2D candidates"))
                    return -1;
                cand = Vc(start);
            }
            ++mgs_fallbacks;
            accepted = 0;
            std::vector<double> n2(1);
            for (int c = 0; c < nc; ++c) {
                const int src = start + c, dst = start + accepted;
                if (dst != src &&
                    !ok(hipMemcpyAsync(Vc(dst), Vc(src), static_cast<size_t>(npw) * 16, hipMemcpyDeviceToDevice, st),
                        "D2D MGS move"))
                    return -1;
                const int nex = dst - start;
                if (nex > 0) {
                    if (!gram(nex, 1, cand, Vc(dst), C.get(), nex)) return -1;
                    if (!gemm(0, npw, 1, nex, -1.0, cand, npw, C.get(), nex, 1.0, Vc(dst), npw)) return -1;
                }
                if (!ok(launch_column_norm2(npw, 1, Vc(dst), npw, norms_d.as<double>(), st), "norm") ||
                    (gamma && !ok(launch_gamma_norm_fix(1, Vc(dst), npw, norms_d.as<double>(), st), "Gamma norm")) ||
                    !ok(pin().d2h(n2.data(), norms_d.get(), 8, st), "D2H norm"))
                    return -1;
                if (n2[0] > kNormThreshold) {
                    const double inv = 1.0 / std::sqrt(n2[0]);
                    if (!ok(pin().h2d(eps_d.get(), &inv, 8, st), "H2D scale") ||
                        !ok(launch_axpby(npw, 1, eps_d.as<double>(), Vc(dst), npw, 0.0, Vc(dst), npw, st), "scale") ||
                        !sync())
                        return -1;
                    ++accepted;
                }
            }
        }
        if (!stage_sync()) return -1;
        stage[1] += now_ms() - t0;
        return accepted;
    }

    // V[:, :ncols] <- V[:, :n] X[:, :ncols], same for HV (serial.hpp:825-945); X = X_d (ld n).
    // The products go to the other buffers of the ping-pong, which then become the basis.
    bool rotate(int n, int ncols) {
        if (!gemm(0, npw, ncols, n, 1.0, vbase, npw, X_d.get(), n, 0.0, valt, npw) ||
            !gemm(0, npw, ncols, n, 1.0, hvbase, npw, X_d.get(), n, 0.0, hvalt, npw))
            return false;
        std::swap(vbase, valt);
        std::swap(hvbase, hvalt);
        if (!g0_real(vbase, ncols) || !g0_real(hvbase, ncols)) return false;  // serial.hpp:939-943, 982-983
        return stage_sync();
    }

    // Collapse to the nb lowest Ritz vectors (serial.hpp:825-945).  w_out (optional) receives the
    // Ritz values: the rebuilt nb x nb subspace is diag(w) to rounding, with eigenvectors I.
    int collapse(int n, bool have_matrices, std::vector<double>* w_out = nullptr) {
        const double t0 = now_ms();
        std::vector<double> w;
        int info = -2;
        if (have_matrices) info = diag(n, w, nb);
        if (info == -1) return -1;
        if (info != 0) {
            if (!update(0, n, n)) return -1;  // build_subspace_matrices_and_diagonalize
            info = diag(n, w, nb);
            if (info == -1) return -1;
            if (info != 0) {
                status = Q3X_ERR_NUMERIC;
                set_error("Davidson: subspace eigensolver failed on a rebuilt subspace");
                return -1;
            }
        }
        if (!rotate(n, nb)) return -1;
        if (!update(0, nb, nb)) return -1;
        if (w_out) *w_out = w;
        stage[5] += now_ms() - t0;
        return nb;
    }

    // The eigenpairs of the subspace right after a collapse: the collapse's Ritz values and the
    // identity (the basis is made of the Ritz vectors), in place of a solve of the rebuilt
    // subspace, whose result differs from these at rounding level.
    bool diag_after_collapse(const std::vector<double>& wc, std::vector<double>& w) {
        w.assign(std::max<size_t>(wc.size(), static_cast<size_t>(mmax)), 0.0);
        std::copy(wc.begin(), wc.end(), w.begin());
        return ok(launch_identity(nb, X_d.as<double>(), nb, st), "identity eigenvectors");
    }
};

}  // namespace
}  // namespace q3x

using namespace q3x;

extern "C" void q3x_davidson_default_options(q3x_davidson_options* o) {
    if (!o) return;
    o->tolerance_ry = 1e-2;  // initial_eigenvalue_tolerance (tolerances.hpp:224)
    o->max_iterations = 20;  // davidson.hpp:1156
    o->basis_factor = 2;     // davidson.hpp:864
    o->n_occupied = 0;
    o->precision = Q3X_FP64;
    o->seed_from_psi = 0;
    o->final_fp64_sweep = 1;
    o->stored_eigenvalues = 0;
}

namespace {
// psi_on_device: psi_inout is a device buffer (npw x nb, ld npw) that stays in HBM
q3x_status davidson_solve(q3x_hamiltonian h, int nb, const q3x_davidson_options* opt, double* psi_inout,
                          bool psi_on_device, double* eig_out, q3x_davidson_report* report) {
    if (!h || !opt || !psi_inout || !eig_out) {
        set_error("q3x_davidson_solve: invalid argument");
        return Q3X_ERR_ARGUMENT;
    }
    if (!h->kp || !h->have_potential) {
        set_error("q3x_davidson_solve: set_potential() and prepare_kpoint() must precede solve");
        return Q3X_ERR_STATE;
    }
    if (nb <= 0 || nb > h->npw || opt->basis_factor < 1 || opt->max_iterations < 0 ||
        (opt->precision != Q3X_FP64 && opt->precision != Q3X_MIXED)) {
        set_error("q3x_davidson_solve: need 0 < nb <= npw, basis_factor >= 1, a valid precision");
        return Q3X_ERR_ARGUMENT;
    }
    Solver s(h);
    s.timing = h->ctx->prof.on;
    s.npw = h->npw;
    s.nb = nb;
    s.nocc = std::max(0, std::min(opt->n_occupied, nb));
    s.mmax = std::max(std::min(opt->basis_factor * nb, s.npw), nb);
    s.precision = opt->precision;
    s.gamma = h->gamma;
    const int npw = s.npw, mmax = s.mmax;
    const double t_start = s.now_ms();
    const size_t col = static_cast<size_t>(npw) * 16;
    const size_t mm16 = static_cast<size_t>(mmax) * mmax * 16;
    if (!s.V.ensure(col * mmax) || !s.HV.ensure(col * mmax) || !s.ws.V2.ensure(col * mmax) ||
        !s.ws.HV2.ensure(col * mmax) || !s.R.ensure(col * std::max(nb, 1)) ||
        !s.C.ensure(mm16) || !s.G.ensure(mm16) || !s.Hs_d.ensure(mm16) || !s.Ss_d.ensure(mm16) ||
        !s.X_d.ensure(mm16) || !s.Xc_d.ensure(static_cast<size_t>(mmax) * nb * 32) ||
        !s.eps_d.ensure(static_cast<size_t>(mmax) * 8) || !s.negeps_d.ensure(static_cast<size_t>(mmax) * 20) ||
        !s.norms_d.ensure(static_cast<size_t>(mmax) * 8) || !s.idx_d.ensure(static_cast<size_t>(mmax) * 4) ||
        !s.work.ensure(s.work_bytes) || !s.eigw.ensure(subspace_eig_workspace(mmax, mmax)))
        return Q3X_ERR_DEVICE;
    s.vbase = s.V.as<double>();
    s.hvbase = s.HV.as<double>();
    s.valt = s.ws.V2.as<double>();
    s.hvalt = s.ws.HV2.as<double>();
    int prec = opt->precision;
    s.fp32_products = prec == Q3X_MIXED && opt->tolerance_ry >= kMixedProductsMinTol;

    // ---- seed (davidson.hpp:1000-1101)
    double t0 = s.now_ms();
    int basis = 0;
    if (opt->seed_from_psi) {
        if (!s.ok(hipMemcpyAsync(s.Vc(0), psi_inout, col * nb,
                                 psi_on_device ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice, s.st),
                  "copy psi in"))
            return s.status;
        const int acc = s.orthogonalize(0, nb);
        if (acc < 0) return s.status;
        basis = acc;
        if (!s.apply(0, basis, prec)) return s.status;
    }
    if (basis < nb) {
        const KPointData& K = h->kp->d;
        std::vector<std::pair<double, int>> ke(npw);
        for (int g = 0; g < npw; ++g) ke[g] = {K.ekin[g], g};
        std::sort(ke.begin(), ke.end());
        const int start = basis, nc = nb - basis;
        std::vector<int> idx(nc);
        for (int c = 0; c < nc; ++c) idx[c] = ke[start + c].second;
        if (!s.ok(launch_zero(s.Vc(start), 2L * npw * nc, s.st), "zero seed") ||
            !s.ok(s.pin().h2d(s.idx_d.get(), idx.data(), nc * 4, s.st), "H2D seed idx") ||
            !s.ok(launch_unit_columns(s.idx_d.as<int>(), nc, s.Vc(start), npw, s.st), "unit columns"))
            return s.status;
        const int acc = s.orthogonalize(start, nc);
        if (acc < 0) return s.status;
        basis += acc;
        if (!s.apply(start, basis - start, prec)) return s.status;
    }
    if (!s.update(0, basis, basis)) return s.status;
    std::vector<double> prev(nb, 0.0), w;
    {  // starting values: the Rayleigh quotients on the diagonal of H_sub
        std::vector<zc> diag(basis);
        char* dst = s.pin().take(static_cast<size_t>(basis) * 16, s.st);
        if (!dst) {
            s.ok(hipErrorOutOfMemory, "pinned staging");
            return s.status;
        }
        if (!s.ok(hipMemcpy2DAsync(dst, 16, s.Hs_d.get(), static_cast<size_t>(mmax + 1) * 16, 16, basis,
                                   hipMemcpyDeviceToHost, s.st),
                  "D2H diagonal") ||
            !s.sync())
            return s.status;
        std::memcpy(diag.data(), dst, static_cast<size_t>(basis) * 16);
        s.pin().rewind();
        for (int b = 0; b < nb && b < basis; ++b) prev[b] = diag[b].real();
    }
    // stored eigenvalues of the k-point state (e.g. carried over from the previous k-point of a
    // band path) replace the diagonal when any of them is nonzero (davidson.hpp:473-493)
    if (opt->stored_eigenvalues && std::any_of(eig_out, eig_out + nb, [](double e) { return e != 0.0; }))
        for (int b = 0; b < nb; ++b) prev[b] = eig_out[b];
    s.stage[0] += s.now_ms() - t0;

    const double tol = opt->tolerance_ry;
    const double empty_tol = std::max(kEmptyFactor * tol, kEmptyFloor);
    int it = 0, unconv = nb, iterations_run = 0, refinements = 0;
    bool done = false;
    bool collapsed = false;  // the subspace was just rebuilt from Ritz vectors with values wc
    std::vector<double> wc;
    int max_iter = opt->max_iterations;
    for (;;) {
        for (it = 0; it < max_iter; ++it) {
            ++iterations_run;
            int info = 0;
            const bool x_identity = collapsed && basis == nb;  // Ritz vectors = basis columns
            if (x_identity) {
                if (!s.diag_after_collapse(wc, w)) return s.status;
            } else {
                info = s.diag(basis, w, nb);
            }
            collapsed = false;
            if (info < 0) return s.status;
            if (info != 0 && basis > nb) {  // davidson.hpp:520-590
                ++s.recovery_collapses;
                basis = s.collapse(basis, false);
                if (basis < 0) return s.status;
                info = s.diag(basis, w, nb);
                if (info < 0) return s.status;
                if (info != 0) {
                    ++s.previous_eigenvalue_keeps;
                    for (int b = 0; b < nb; ++b) w[b] = prev[b];
                    // the collapsed basis is made of Ritz vectors: their coefficients are the
                    // identity (the reference goes on with whatever the failed solve left)
                    if (!s.ok(launch_identity(nb, s.X_d.as<double>(), nb, s.st), "identity eigenvectors"))
                        return s.status;
                }
            }
            t0 = s.now_ms();
            std::vector<int> src;
            unconv = 0;
            for (int b = 0; b < nb; ++b) {
                const double ch = std::abs(w[b] - prev[b]);
                const double bt = b < s.nocc ? tol : empty_tol;
                const bool c = (it != 0 && ch < bt);
                if (!c) {
                    ++unconv;
                    if (static_cast<int>(src.size()) < mmax - basis) src.push_back(b);
                }
            }
            if (unconv == 0) {
                done = true;
                s.stage[6] += s.now_ms() - t0;
                break;
            }
            // corrections into V[:, basis:basis+nc]: (HV - e V) X_c, preconditioned
            const int nc = static_cast<int>(src.size());
            if (nc > 0) {
                std::vector<double> ec(nc), nec(nc);
                for (int c = 0; c < nc; ++c) {
                    ec[c] = w[src[c]];
                    nec[c] = -ec[c];
                }
                double* corr = s.R.as<double>();  // formed in R, orthonormalised into V[:, basis:]
                double* xc = s.Xc_d.as<double>();
                // e, -e and the band indices of the corrections in one upload: [e | -e | idx]
                std::vector<char> packed(static_cast<size_t>(nc) * 20);
                std::memcpy(packed.data(), ec.data(), static_cast<size_t>(nc) * 8);
                std::memcpy(packed.data() + static_cast<size_t>(nc) * 8, nec.data(), static_cast<size_t>(nc) * 8);
                std::memcpy(packed.data() + static_cast<size_t>(nc) * 16, src.data(), static_cast<size_t>(nc) * 4);
                double* eps_c = s.negeps_d.as<double>();
                double* negeps_c = eps_c + nc;
                const int* idx_c = reinterpret_cast<const int*>(eps_c + 2L * nc);
                if (!s.ok(s.pin().h2d(s.negeps_d.get(), packed.data(), packed.size(), s.st), "H2D e, -e, idx"))
                    return s.status;
                if (x_identity) {
                    if (!s.ok(launch_residual_columns_teter(npw, nc, idx_c, eps_c, h->ekin.as<double>(), s.Vc(0), npw,
                                                            s.HVc(0), npw, corr, npw, s.st),
                              "corrections"))
                        return s.status;
                } else if (!s.ok(launch_gather_columns(basis, nc, idx_c, nullptr, s.X_d.as<double>(), basis, xc, basis,
                                                       s.st),
                                 "gather X") ||
                           !s.ok(launch_gather_columns(basis, nc, idx_c, negeps_c, s.X_d.as<double>(), basis,
                                                       xc + 2L * basis * nc, basis, s.st),
                                 "gather -e X") ||
                           !s.gemm(0, npw, nc, basis, 1.0, s.Vc(0), npw, xc + 2L * basis * nc, basis, 0.0, corr,
                                   npw) ||
                           !s.gemm(0, npw, nc, basis, 1.0, s.HVc(0), npw, xc, basis, 1.0, corr, npw) ||
                           !s.ok(launch_teter(npw, nc, h->ekin.as<double>(), eps_c, corr, npw, s.st), "Teter")) {
                    return s.status;
                }
                if (!s.g0_real(corr, nc)) return s.status;  // serial.hpp:1129-1130
            }
            if (!s.stage_sync()) return s.status;
            s.stage[6] += s.now_ms() - t0;
            const int accepted = s.orthogonalize(basis, nc, true);
            if (accepted < 0) return s.status;
            if (!s.apply(basis, accepted, prec)) return s.status;
            const int old = basis;
            basis += accepted;
            if (!s.update(old, accepted, basis)) return s.status;
            if (basis >= mmax) {
                basis = s.collapse(basis, true, &wc);
                if (basis < 0) return s.status;
                collapsed = static_cast<int>(wc.size()) >= nb;
            }
            for (int b = 0; b < nb; ++b) prev[b] = w[b];
        }
        if (!done) {  // davidson.hpp:812-827: final subspace solve after the cap
            const int info = s.diag(basis, w, nb);
            if (info < 0) return s.status;
            if (info != 0) {
                set_error("Davidson: final subspace eigensolve failed");
                return Q3X_ERR_NUMERIC;
            }
        }
        if (prec == Q3X_MIXED && opt->final_fp64_sweep && refinements == 0) {
            // fp64 polish: Ritz vectors of the mixed solve, H re-applied in fp64, iterate on
            ++refinements;
            if (!s.rotate(basis, nb)) return s.status;
            basis = nb;
            prec = Q3X_FP64;
            s.fp32_products = false;
            if (!s.apply(0, nb, prec)) return s.status;
            if (!s.update(0, nb, nb)) return s.status;
            for (int b = 0; b < nb; ++b) prev[b] = w[b];
            done = false;
            max_iter = std::max(opt->max_iterations, 1);
            continue;
        }
        break;
    }
    // extract (serial.hpp:947-993): psi = V X[:, :nb]; then the fp64 residual sweep
    // r_n = H psi_n - e_n psi_n, reported as max_n ||r_n||
    if (!s.rotate(basis, nb)) return s.status;
    double maxres = 0.0;
    {
        std::vector<double> e(nb), n2(nb);
        for (int b = 0; b < nb; ++b) e[b] = -w[b];
        double* res = s.R.as<double>();
        if (!s.ok(s.pin().h2d(s.eps_d.get(), e.data(), nb * 8, s.st), "H2D e") ||
            !s.ok(launch_axpby(npw, nb, s.eps_d.as<double>(), s.Vc(0), npw, 0.0, res, npw, s.st), "r = -e psi") ||
            !s.ok(launch_axpby(npw, nb, nullptr, s.HVc(0), npw, 1.0, res, npw, s.st), "r += H psi") ||
            !s.ok(launch_column_norm2(npw, nb, res, npw, s.norms_d.as<double>(), s.st), "norms") ||
            (s.gamma && !s.ok(launch_gamma_norm_fix(nb, res, npw, s.norms_d.as<double>(), s.st), "Gamma norms")) ||
            !s.ok(s.pin().d2h(n2.data(), s.norms_d.get(), nb * 8, s.st), "D2H norms"))
            return s.status;
        for (int b = 0; b < nb; ++b) maxres = std::max(maxres, std::sqrt(n2[b]));
    }
    if (!s.ok(hipMemcpyAsync(psi_inout, s.Vc(0), col * nb,
                             psi_on_device ? hipMemcpyDeviceToDevice : hipMemcpyDeviceToHost, s.st),
              "copy psi out") ||
        !s.sync())
        return s.status;
    for (int b = 0; b < nb; ++b) eig_out[b] = w[b];
    s.stage[7] = s.now_ms() - t_start;
    if (report) {
        report->iterations = iterations_run;
        report->unconverged = done ? 0 : unconv;
        report->h_applications = s.h_applications;
        report->fp64_refinements = refinements;
        report->max_residual_norm = maxres;
        report->eig_barrier_redos = s.eig_stats.barrier_redos;
        report->mgs_fallbacks = s.mgs_fallbacks;
        report->recovery_collapses = s.recovery_collapses;
        report->previous_eigenvalue_keeps = s.previous_eigenvalue_keeps;
        for (int i = 0; i < 8; ++i) report->stage_ms[i] = s.stage[i];
    }
    return Q3X_OK;
}
}  // namespace

extern "C" q3x_status q3x_davidson_solve(q3x_hamiltonian h, int nb, const q3x_davidson_options* opt,
                                         double* psi_inout, double* eig_out, q3x_davidson_report* report) {
    try {
        return davidson_solve(h, nb, opt, psi_inout, false, eig_out, report);
    } catch (const std::exception& e) {  // nothing crosses the C ABI as an exception
        set_error(std::string("q3x_davidson_solve: ") + e.what());
        return Q3X_ERR_INTERNAL;
    }
}

extern "C" q3x_status q3x_davidson_solve_device(q3x_hamiltonian h, int nb, const q3x_davidson_options* opt,
                                                void* d_psi_inout, double* eig_out, q3x_davidson_report* report) {
    try {
        return davidson_solve(h, nb, opt, static_cast<double*>(d_psi_inout), true, eig_out, report);
    } catch (const std::exception& e) {  // nothing crosses the C ABI as an exception
        set_error(std::string("q3x_davidson_solve_device: ") + e.what());
        return Q3X_ERR_INTERNAL;
    }
}
