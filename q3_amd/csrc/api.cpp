// C ABI: device context, device buffers, Hamiltonian applicator.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <stdexcept>
#include <string>
#include <vector>

#include "aux_kernels.h"
#include "gamma_kernels.h"
#include "potential.h"
#include "runtime.h"
#include "subspace_eig.h"
#include "zgemm.h"

namespace q3x {

namespace {
thread_local std::string g_last_error;
}

void set_error(const std::string& msg) { g_last_error = msg; }

bool hip_ok(hipError_t e, const char* what) {
    if (e == hipSuccess) return true;
    set_error(std::string(what) + ": " + hipGetErrorString(e));
    return false;
}

bool DevBuf::ensure(size_t bytes) {
    if (bytes <= bytes_ && p_) return true;
    release();
    if (bytes == 0) bytes = 16;
    if (!hip_ok(hipMalloc(&p_, bytes), "hipMalloc")) {
        p_ = nullptr;
        bytes_ = 0;
        return false;
    }
    bytes_ = bytes;
    return true;
}

void DevBuf::release() {
    if (p_) (void)hipFree(p_);
    p_ = nullptr;
    bytes_ = 0;
}

bool DevBuf::upload(const void* src, size_t bytes, hipStream_t st, HostStage* pin) {
    if (!ensure(bytes)) return false;
    if (bytes == 0) return true;
    if (pin && bytes <= (size_t(16) << 20)) return hip_ok(pin->h2d(p_, src, bytes, st), "pinned H2D");
    return hip_ok(hipMemcpyAsync(p_, src, bytes, hipMemcpyHostToDevice, st), "hipMemcpyAsync H2D") &&
           hip_ok(hipStreamSynchronize(st), "hipStreamSynchronize");
}

HostStage::~HostStage() {
    if (p_) (void)hipHostFree(p_);
}

char* HostStage::take(size_t bytes, hipStream_t st) {
    const size_t n = (std::max<size_t>(bytes, 1) + 255) & ~static_cast<size_t>(255);
    if (off_ + n > cap_) {
        if (hipStreamSynchronize(st) != hipSuccess) return nullptr;
        off_ = 0;
        if (n > cap_) {
            const size_t want = std::max<size_t>(n, std::max<size_t>(2 * cap_, size_t(4) << 20));
            if (p_) (void)hipHostFree(p_);
            p_ = nullptr;
            cap_ = 0;
            void* q = nullptr;
            if (hipHostMalloc(&q, want, hipHostMallocDefault) != hipSuccess) return nullptr;
            p_ = static_cast<char*>(q);
            cap_ = want;
        }
    }
    char* r = p_ + off_;
    off_ += n;
    return r;
}

hipError_t HostStage::h2d(void* dst, const void* src, size_t bytes, hipStream_t st) {
    if (bytes == 0) return hipSuccess;
    char* s = take(bytes, st);
    if (!s) return hipErrorOutOfMemory;
    std::memcpy(s, src, bytes);
    return hipMemcpyAsync(dst, s, bytes, hipMemcpyHostToDevice, st);
}

hipError_t HostStage::d2h(void* dst, const void* src, size_t bytes, hipStream_t st) {
    if (bytes == 0) return hipSuccess;
    char* s = take(bytes, st);
    if (!s) return hipErrorOutOfMemory;
    hipError_t e = hipMemcpyAsync(s, src, bytes, hipMemcpyDeviceToHost, st);
    if (e == hipSuccess) e = hipStreamSynchronize(st);
    if (e != hipSuccess) return e;
    std::memcpy(dst, s, bytes);
    off_ = 0;
    return hipSuccess;
}

int Profiler::mark(hipStream_t st) {
    if (!on) return -1;
    if (used == pool.size()) {
        hipEvent_t e;
        if (hipEventCreate(&e) != hipSuccess) return -1;
        pool.push_back(e);
    }
    (void)hipEventRecord(pool[used], st);
    return static_cast<int>(used++);
}

bool Profiler::flush(hipStream_t st) {
    if (!hip_ok(hipStreamSynchronize(st), "profiler sync")) return false;
    for (const Span& s : spans) {
        float ms = 0.f;
        if (hipEventElapsedTime(&ms, pool[s.e0], pool[s.e1]) == hipSuccess) {
            total_ms[s.id] += ms;
            count[s.id] += 1;
        }
    }
    spans.clear();
    used = 0;
    return true;
}

Profiler::~Profiler() {
    for (hipEvent_t e : pool) (void)hipEventDestroy(e);
}

namespace {

// Radix plan for one FFT length.  Every transform is run by one wave64, so each stage must
// have n / R <= 64 butterflies per transform.  Search all orderings of supported radices for
// the fewest stages; ties go to the earlier radices of `pref`.  Radices stop at 8: larger
// composite butterflies need more registers than an FP64 stage can hold without spilling
// (96 = 8 * 6 * 2, 48 = 8 * 6, 30 = 6 * 5).
bool plan_search(int n, int rem, int depth, int maxdepth, int* cur, FftPlan& best) {
    static const int pref[] = {8, 6, 5, 4, 3, 2};
    if (rem == 1) {
        best.nstage = depth;
        for (int i = 0; i < depth; ++i) best.radix[i] = cur[i];
        return true;
    }
    if (depth == maxdepth) return false;
    for (int q : pref)
        if (rem % q == 0 && n / q <= 64) {
            cur[depth] = q;
            if (plan_search(n, rem / q, depth + 1, maxdepth, cur, best)) return true;
        }
    return false;
}

FftPlan make_plan(int n) {
    FftPlan p{};
    p.n = n;
    p.nstage = -1;
    int cur[8];
    for (int maxd = 1; maxd <= 8 && n >= 2 && n <= 1024; ++maxd)
        if (plan_search(n, n, 0, maxd, cur, p)) break;
    return p;
}

bool build_tables(FftTables& t, int n, hipStream_t st, HostStage* pin) {
    t.plan = make_plan(n);
    if (t.plan.nstage < 0) {
        set_error("FFT length is not 2,3,5-smooth");
        return false;
    }
    // stage-major twiddles: stage s (radix R, stride ns) holds [r-1][jn] = exp(-2 pi i r jn / (ns R))
    std::vector<double> d(2 * n, 0.0);
    std::vector<float> f(2 * n, 0.0f);
    const long double two_pi = 6.283185307179586476925286766559005768L;
    int off = 0, ns = 1;
    for (int s = 0; s < t.plan.nstage; ++s) {
        const int R = t.plan.radix[s];
        t.plan.twoff[s] = off;
        for (int r = 1; r < R; ++r)
            for (int jn = 0; jn < ns; ++jn, ++off) {
                const long k = static_cast<long>(r) * jn * (n / (ns * R));
                const long double a = -two_pi * static_cast<long double>(k % n) / n;
                d[2 * off] = static_cast<double>(cosl(a));
                d[2 * off + 1] = static_cast<double>(sinl(a));
                f[2 * off] = static_cast<float>(d[2 * off]);
                f[2 * off + 1] = static_cast<float>(d[2 * off + 1]);
            }
        ns *= R;
    }
    std::vector<double> pd(2 * n);
    std::vector<float> pf(2 * n);
    for (int j = 0; j < n; ++j) {
        const long double a = -two_pi * static_cast<long double>(j) / n;
        pd[2 * j] = static_cast<double>(cosl(a));
        pd[2 * j + 1] = static_cast<double>(sinl(a));
        pf[2 * j] = static_cast<float>(pd[2 * j]);
        pf[2 * j + 1] = static_cast<float>(pd[2 * j + 1]);
    }
    return t.tw64.upload(d.data(), d.size() * sizeof(double), st, pin) &&
           t.tw32.upload(f.data(), f.size() * sizeof(float), st, pin) &&
           t.plain64.upload(pd.data(), pd.size() * sizeof(double), st, pin) &&
           t.plain32.upload(pf.data(), pf.size() * sizeof(float), st, pin);
}

HLocalArgs make_args(const q3x_hamiltonian_s* h, bool fp32) {
    HLocalArgs a{};
    a.p1 = h->fft[0].plan;
    a.p2 = h->fft[1].plan;
    a.p3 = h->fft[2].plan;
    a.tw1 = fp32 ? h->fft[0].tw32.get() : h->fft[0].tw64.get();
    a.tw2 = fp32 ? h->fft[1].tw32.get() : h->fft[1].tw64.get();
    a.tw3 = fp32 ? h->fft[2].tw32.get() : h->fft[2].tw64.get();
    a.twp = fp32 ? h->fft[0].plain32.get() : h->fft[0].plain64.get();
    a.twp3 = fp32 ? h->fft[2].plain32.get() : h->fft[2].plain64.get();
    a.nst = h->nst;
    a.ncol = h->ncol;
    a.sticks_per_block = h->sticks_per_block;
    a.rows_per_pass = fp32 ? h->rows_per_pass32 : h->rows_per_pass64;
    a.stick_off = h->stick_off.as<int>();
    a.stick_i2 = h->stick_i2.as<int>();
    a.stick_col = h->stick_col.as<int>();
    a.col_i1 = h->col_i1.as<int>();
    a.ent_pw = h->ent_pw.as<int>();
    a.ent_pos = h->ent_pos.as<int>();
    a.ekin = h->ekin.as<double>();
    a.outside = h->outside.as<int>();
    a.n_outside = h->n_outside;
    return a;
}

// Largest band chunk whose stick buffer (and plane slots, for planes kept outside LDS) stays under 2 GiB.
int band_chunk(const q3x_hamiltonian_s* h, const PlaneTlConfig& tl, int nvec) {
    size_t per = static_cast<size_t>(h->dims[2]) * h->nst * 16;
    if (tl.ok && tl.plane_global) per += static_cast<size_t>(h->dims[2]) * tl.plane_bytes;
    const size_t cap = size_t(2) << 30;
    int c = static_cast<int>(std::max<size_t>(1, cap / std::max<size_t>(per, 1)));
    return std::min(c, nvec);
}

}  // namespace

q3x_status hamiltonian_apply_device(q3x_hamiltonian h, const double* d_psi, long ldpsi, double* d_hpsi, long ldh,
                                    int nvec, int precision) {
    if (!h || !h->kp) {
        set_error("apply: prepare_kpoint() must be called before apply (serial.hpp:487-492)");
        return Q3X_ERR_STATE;
    }
    if (!h->have_potential) {
        set_error("apply: set_potential() must be called before apply");
        return Q3X_ERR_STATE;
    }
    if (nvec <= 0) return Q3X_OK;
    if (precision != Q3X_FP64 && precision != Q3X_MIXED) {
        set_error("apply: precision must be Q3X_FP64 or Q3X_MIXED");
        return Q3X_ERR_ARGUMENT;
    }
    const bool fp32 = precision == Q3X_MIXED;
    hipStream_t st = h->ctx->stream;
    const HLocalArgs args = make_args(h, fp32);
    const PlaneTlConfig& tl = fp32 ? h->tl32 : h->tl64;
    // Gamma trick: the transforms run on packed pairs of bands (nvec_t columns of the extended list)
    const int nvec_t = h->gamma ? (nvec + 1) / 2 : nvec;
    const int chunk = band_chunk(h, tl, nvec_t);
    const size_t cs = fp32 ? 8 : 16;
    if (!h->zbuf.ensure(static_cast<size_t>(chunk) * h->dims[2] * h->nst * cs)) return Q3X_ERR_DEVICE;
    if (tl.ok && tl.plane_global && !h->plane_slots.ensure(static_cast<size_t>(chunk) * h->dims[2] * tl.plane_bytes))
        return Q3X_ERR_DEVICE;
    const void* vplane = fp32 ? h->v32.get() : h->v64.get();
    Profiler& pr = h->ctx->prof;
    if (h->gamma) {
        const size_t gb = static_cast<size_t>(h->npw_ext) * chunk * 16;
        if (!h->gpack.ensure(gb) || !h->gout.ensure(gb)) return Q3X_ERR_DEVICE;
    }
    for (int b0 = 0; b0 < nvec_t; b0 += chunk) {
        const int nb = std::min(chunk, nvec_t - b0);
        hipEvent_t marks[4];
        int mi[4] = {-1, -1, -1, -1};
        if (pr.on) {
            for (int i = 0; i < 4; ++i) {
                if (pr.used == pr.pool.size()) {
                    hipEvent_t e;
                    if (!hip_ok(hipEventCreate(&e), "hipEventCreate")) return Q3X_ERR_DEVICE;
                    pr.pool.push_back(e);
                }
                mi[i] = static_cast<int>(pr.used++);
                marks[i] = pr.pool[mi[i]];
            }
        }
        const double* in = d_psi + 2L * b0 * ldpsi;
        double* out = d_hpsi + 2L * b0 * ldh;
        long ldin = ldpsi, ldout = ldh;
        int nbands = 0;
        if (h->gamma) {  // bands 2 b0 .. of this chunk packed in pairs (serial.hpp:283-418)
            nbands = std::min(2 * nb, nvec - 2 * b0);
            if (!hip_ok(launch_gamma_pack(h->npw, nbands, d_psi + 4L * b0 * ldpsi, ldpsi, h->gpack.as<double>(),
                                          h->npw_ext, st),
                        "Gamma pack"))
                return Q3X_ERR_DEVICE;
            in = h->gpack.as<double>();
            out = h->gout.as<double>();
            ldin = ldout = h->npw_ext;
        }
        if (!hip_ok(launch_hlocal(args, fp32, vplane, h->zbuf.get(), in, ldin, out, ldout, nb, st,
                                  pr.on ? marks : nullptr, &tl, h->plane_slots.get(), fp32 ? &h->tz32 : &h->tz64),
                    "H_local kernels"))
            return Q3X_ERR_DEVICE;
        if (h->gamma &&
            !hip_ok(launch_gamma_unpack(h->npw, nbands, h->gout.as<double>(), h->npw_ext, d_hpsi + 4L * b0 * ldh, ldh,
                                        st),
                    "Gamma unpack"))
            return Q3X_ERR_DEVICE;
        pr.span(K_ZINV, mi[0], mi[1]);
        pr.span(K_PLANE, mi[1], mi[2]);
        pr.span(K_ZFWD, mi[2], mi[3]);
    }
    if (h->have_projectors && h->nproj > 0) {
        const int np = h->nproj, npw = h->npw;
        if (!h->P.ensure(static_cast<size_t>(np) * nvec * cs) || !h->W.ensure(static_cast<size_t>(np) * nvec * cs))
            return Q3X_ERR_DEVICE;
        const size_t ws = static_cast<size_t>(64) << 20;
        if (!h->gemm_work.ensure(ws)) return Q3X_ERR_DEVICE;
        // P = beta^H psi (serial.hpp:572-584)
        const int m0 = pr.mark(st);
        hipError_t e = fp32 ? zgemm(GemmTypes::F_D_F_F, 1, np, nvec, npw, 1.0, h->beta32.get(), npw, d_psi, ldpsi,
                                    0.0, h->P.get(), np, h->gemm_work.get(), ws, st)
                            : zgemm(GemmTypes::D_D_D_D, 1, np, nvec, npw, 1.0, h->beta64.get(), npw, d_psi, ldpsi,
                                    0.0, h->P.get(), np, h->gemm_work.get(), ws, st);
        if (!hip_ok(e, "projector overlap GEMM")) return Q3X_ERR_DEVICE;
        // Gamma: 2 Re(beta^H psi) less the G = 0 term (serial.hpp:519-556)
        if (h->gamma && !hip_ok(launch_gamma_overlap_fix(fp32, np, nvec, fp32 ? h->beta32.get() : h->beta64.get(), npw,
                                                         d_psi, ldpsi, h->P.get(), np, st),
                                "Gamma projector overlaps"))
            return Q3X_ERR_DEVICE;
        const int m1 = pr.mark(st);
        pr.span(K_KB_OVERLAP, m0, m1);
        // W = D P per atom block (serial.hpp:587-679)
        if (!hip_ok(launch_dblock_apply(fp32, np, nvec, h->pstart.as<int>(), h->psize.as<int>(), h->poff.as<long>(),
                                        h->dblocks.as<double>(), h->P.get(), h->W.get(), st),
                    "D block kernel"))
            return Q3X_ERR_DEVICE;
        const int m2 = pr.mark(st);
        pr.span(K_KB_DBLOCK, m1, m2);
        // Hpsi += beta W (serial.hpp:688-700)
        e = fp32 ? zgemm(GemmTypes::F_F_F_D, 0, npw, nvec, np, 1.0, h->beta32.get(), npw, h->W.get(), np, 1.0, d_hpsi,
                         ldh, nullptr, 0, st)
                 : zgemm(GemmTypes::D_D_D_D, 0, npw, nvec, np, 1.0, h->beta64.get(), npw, h->W.get(), np, 1.0, d_hpsi,
                         ldh, nullptr, 0, st);
        if (!hip_ok(e, "projector update GEMM")) return Q3X_ERR_DEVICE;
        const int m3 = pr.mark(st);
        pr.span(K_KB_UPDATE, m2, m3);
    }
    return Q3X_OK;
}

namespace {
size_t align256(size_t b) { return (b + 255) & ~static_cast<size_t>(255); }

// Scratch bytes of one band chunk of nb bands carved from a fixed region: stick buffer, plane
// slots (planes kept outside LDS), P and W of the projector chain, each 256-byte aligned.
size_t fixed_chunk_bytes(const q3x_hamiltonian_s* h, const PlaneTlConfig& tl, int nb, size_t cs) {
    size_t b = align256(static_cast<size_t>(nb) * h->dims[2] * h->nst * cs);
    if (tl.ok && tl.plane_global) b += align256(static_cast<size_t>(nb) * h->dims[2] * tl.plane_bytes);
    if (h->have_projectors && h->nproj > 0) b += 2 * align256(static_cast<size_t>(h->nproj) * nb * cs);
    return b;
}
}  // namespace

q3x_status hamiltonian_apply_fixed(q3x_hamiltonian h, const double* d_psi, double* d_hpsi, int nvec, int precision,
                                   char* scratch, size_t scratch_bytes) {
    if (!h || !h->kp) {
        set_error("apply: prepare_kpoint() must be called before apply (serial.hpp:487-492)");
        return Q3X_ERR_STATE;
    }
    if (!h->have_potential) {
        set_error("apply: set_potential() must be called before apply");
        return Q3X_ERR_STATE;
    }
    if (nvec < 0) {
        set_error("apply: negative band count");
        return Q3X_ERR_ARGUMENT;
    }
    if (nvec == 0) return Q3X_OK;
    if (precision != Q3X_FP64 && precision != Q3X_MIXED) {
        set_error("apply: precision must be Q3X_FP64 or Q3X_MIXED");
        return Q3X_ERR_ARGUMENT;
    }
    if (!scratch || !d_psi || !d_hpsi) {
        set_error("apply: null device pointer");
        return Q3X_ERR_ARGUMENT;
    }
    if (h->gamma) {
        set_error("apply (no allocation): not available for a Gamma-trick basis");
        return Q3X_ERR_ARGUMENT;
    }
    const bool fp32 = precision == Q3X_MIXED;
    const size_t cs = fp32 ? 8 : 16;
    const PlaneTlConfig& tl = fp32 ? h->tl32 : h->tl64;
    const size_t one = fixed_chunk_bytes(h, tl, 1, cs);
    if (one > scratch_bytes) {
        set_error("apply (no allocation): one band needs " + std::to_string(one) + " bytes of scratch (stick buffer" +
                  std::string(tl.ok && tl.plane_global ? ", plane slots" : "") + ", projector products), more than the " +
                  std::to_string(scratch_bytes) + " bytes of the fixed scratch buffer");
        return Q3X_ERR_ARGUMENT;
    }
    // largest chunk of bands whose buffers fit, leaving the split-K workspace the rest
    int chunk = static_cast<int>(std::min<size_t>(static_cast<size_t>(nvec), scratch_bytes / one));
    while (chunk > 1 && fixed_chunk_bytes(h, tl, chunk, cs) > scratch_bytes) --chunk;
    const size_t zb = align256(static_cast<size_t>(chunk) * h->dims[2] * h->nst * cs);
    const size_t sb = tl.ok && tl.plane_global ? align256(static_cast<size_t>(chunk) * h->dims[2] * tl.plane_bytes) : 0;
    const bool nonlocal = h->have_projectors && h->nproj > 0;
    const size_t pb = nonlocal ? align256(static_cast<size_t>(h->nproj) * chunk * cs) : 0;
    char* zbuf = scratch;
    char* slots = sb ? scratch + zb : nullptr;
    char* P = scratch + zb + sb;
    char* W = P + pb;
    char* work = W + pb;
    const size_t used = zb + sb + 2 * pb;
    const size_t ws = scratch_bytes > used ? std::min(scratch_bytes - used, static_cast<size_t>(64) << 20) : 0;
    hipStream_t st = h->ctx->stream;
    const HLocalArgs args = make_args(h, fp32);
    const void* vplane = fp32 ? h->v32.get() : h->v64.get();
    const long ld = h->npw;
    for (int b0 = 0; b0 < nvec; b0 += chunk) {
        const int nb = std::min(chunk, nvec - b0);
        const double* psi = d_psi + 2L * b0 * ld;
        double* hpsi = d_hpsi + 2L * b0 * ld;
        if (!hip_ok(launch_hlocal(args, fp32, vplane, zbuf, psi, ld, hpsi, ld, nb, st, nullptr, &tl, slots,
                                  fp32 ? &h->tz32 : &h->tz64),
                    "H_local kernels"))
            return Q3X_ERR_DEVICE;
        if (!nonlocal) continue;
        const int np = h->nproj, npw = h->npw;
        void* wk = ws ? work : nullptr;
        hipError_t e = fp32 ? zgemm(GemmTypes::F_D_F_F, 1, np, nb, npw, 1.0, h->beta32.get(), npw, psi, ld, 0.0, P, np,
                                    wk, ws, st)
                            : zgemm(GemmTypes::D_D_D_D, 1, np, nb, npw, 1.0, h->beta64.get(), npw, psi, ld, 0.0, P, np,
                                    wk, ws, st);
        if (!hip_ok(e, "projector overlap GEMM")) return Q3X_ERR_DEVICE;
        if (!hip_ok(launch_dblock_apply(fp32, np, nb, h->pstart.as<int>(), h->psize.as<int>(), h->poff.as<long>(),
                                        h->dblocks.as<double>(), P, W, st),
                    "D block kernel"))
            return Q3X_ERR_DEVICE;
        e = fp32 ? zgemm(GemmTypes::F_F_F_D, 0, npw, nb, np, 1.0, h->beta32.get(), npw, W, np, 1.0, hpsi, ld, nullptr,
                         0, st)
                 : zgemm(GemmTypes::D_D_D_D, 0, npw, nb, np, 1.0, h->beta64.get(), npw, W, np, 1.0, hpsi, ld, nullptr,
                         0, st);
        if (!hip_ok(e, "projector update GEMM")) return Q3X_ERR_DEVICE;
    }
    return Q3X_OK;
}

}  // namespace q3x

using namespace q3x;

namespace q3x {
double device_bytes_held(const q3x_context_s* c, const q3x_hamiltonian_s* h) {
    double b = 0.0;
    if (c) {
        const DavidsonWorkspace& w = c->dav;
        for (const DevBuf* d : {&c->gemm_work, &c->eig_work, &w.V, &w.HV, &w.V2, &w.HV2, &w.R, &w.C, &w.G, &w.Hs, &w.Ss,
                                &w.X, &w.Xc, &w.eps, &w.negeps, &w.work, &w.norms, &w.idx, &w.eigw})
            b += static_cast<double>(d->size());
    }
    if (h) {
        for (const DevBuf* d : {&h->v64, &h->v32, &h->stick_off, &h->stick_i2, &h->stick_col, &h->col_i1, &h->ent_pw,
                                &h->ent_pos, &h->ekin, &h->outside, &h->beta64, &h->beta32, &h->dblocks, &h->pstart,
                                &h->psize, &h->poff, &h->zbuf, &h->P, &h->W, &h->gemm_work, &h->io_psi, &h->io_hpsi,
                                &h->plane_slots, &h->gpack, &h->gout, &h->kb_g, &h->kb_tables, &h->kb_form,
                                &h->kb_ints, &h->kb_dbl, &h->dens_w})
            b += static_cast<double>(d->size());
        for (const FftTables& t : h->fft)
            b += static_cast<double>(t.tw64.size() + t.tw32.size() + t.plain64.size() + t.plain32.size());
    }
    return b;
}

q3x_status hamiltonian_rebind(q3x_hamiltonian h, q3x_basis b) {
    if (!h || !b) {
        set_error("hamiltonian rebind: invalid argument");
        return Q3X_ERR_ARGUMENT;
    }
    if (!hip_ok(hipStreamSynchronize(h->ctx->stream), "sync")) return Q3X_ERR_DEVICE;
    // the FFT tables depend on the grid sizes only: kept when they do not change
    const bool same_grid = h->fft[0].plan.n == b->d.wave[0] && h->fft[1].plan.n == b->d.wave[1] &&
                           h->fft[2].plan.n == b->d.wave[2] && h->fft[0].plain64.size() > 0;
    h->basis = b;
    for (int d = 0; d < 3; ++d) h->dims[d] = b->d.wave[d];
    h->npoints = static_cast<long>(h->dims[0]) * h->dims[1] * h->dims[2];
    h->have_potential = false;
    h->kp = nullptr;
    h->npw = h->nst = h->ncol = h->n_outside = 0;
    h->gamma = false;
    h->npw_ext = 0;
    h->nproj = 0;
    h->have_projectors = false;
    if (!same_grid)
        for (int d = 0; d < 3; ++d)
            if (!build_tables(h->fft[d], h->dims[d], h->ctx->stream, &h->ctx->pin)) return Q3X_ERR_DEVICE;
    return Q3X_OK;
}
}  // namespace q3x

extern "C" {

const char* q3x_last_error(void) { return g_last_error.c_str(); }
const char* q3x_version(void) { return "q3x 0.1 (gfx950)"; }

q3x_status q3x_context_create(int device, q3x_context* out) {
    try {
        if (!out) {
            set_error("q3x_context_create: null out");
            return Q3X_ERR_ARGUMENT;
        }
        int n = 0;
        if (!hip_ok(hipGetDeviceCount(&n), "hipGetDeviceCount")) return Q3X_ERR_DEVICE;
        if (device < 0 || device >= n) {
            set_error("q3x_context_create: no such device");
            return Q3X_ERR_DEVICE;
        }
        if (!hip_ok(hipSetDevice(device), "hipSetDevice")) return Q3X_ERR_DEVICE;
        // device address of the no-allocation scratch array, once per device (this one is current)
        if (!noalloc_scratch_init(device)) return Q3X_ERR_DEVICE;
        auto* c = new q3x_context_s;
        c->device = device;
        if (!hip_ok(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking), "hipStreamCreate")) {
            delete c;
            return Q3X_ERR_DEVICE;
        }
        int lds = 0, cus = 0;
        (void)hipDeviceGetAttribute(&lds, hipDeviceAttributeMaxSharedMemoryPerBlock, device);
        (void)hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, device);
        c->max_lds = lds > 0 ? static_cast<size_t>(lds) : 65536;
        c->num_cus = cus;
        *out = c;
        return Q3X_OK;
    } catch (const std::exception& e) {  // nothing crosses the C ABI as an exception
        set_error(std::string("q3x_context_create: ") + e.what());
        return Q3X_ERR_INTERNAL;
    }
}

q3x_status q3x_context_destroy(q3x_context ctx) {
    try {
        if (!ctx) return Q3X_OK;
        (void)hipStreamSynchronize(ctx->stream);
        (void)hipStreamDestroy(ctx->stream);
        delete ctx;
        return Q3X_OK;
    } catch (const std::exception& e) {  // nothing crosses the C ABI as an exception
        set_error(std::string("q3x_context_destroy: ") + e.what());
        return Q3X_ERR_INTERNAL;
    }
}

q3x_status q3x_context_synchronize(q3x_context ctx) {
    try {
        if (!ctx) {
            set_error("null context");
            return Q3X_ERR_ARGUMENT;
        }
        return hip_ok(hipStreamSynchronize(ctx->stream), "hipStreamSynchronize") ? Q3X_OK : Q3X_ERR_DEVICE;
    } catch (const std::exception& e) {  // nothing crosses the C ABI as an exception
        set_error(std::string("q3x_context_synchronize: ") + e.what());
        return Q3X_ERR_INTERNAL;
    }
}

q3x_status q3x_profile_enable(q3x_context ctx, int enable) {
    try {
        if (!ctx) {
            set_error("null context");
            return Q3X_ERR_ARGUMENT;
        }
        if (!ctx->prof.flush(ctx->stream)) return Q3X_ERR_DEVICE;
        ctx->prof.on = enable != 0;
        return Q3X_OK;
    } catch (const std::exception& e) {  // nothing crosses the C ABI as an exception
        set_error(std::string("q3x_profile_enable: ") + e.what());
        return Q3X_ERR_INTERNAL;
    }
}

q3x_status q3x_profile_read(q3x_context ctx, int n, double* total_ms, long* count, int reset) {
    try {
        if (!ctx || n < 0) {
            set_error("q3x_profile_read: invalid argument");
            return Q3X_ERR_ARGUMENT;
        }
        if (!ctx->prof.flush(ctx->stream)) return Q3X_ERR_DEVICE;
        for (int i = 0; i < n && i < K_COUNT; ++i) {
            if (total_ms) total_ms[i] = ctx->prof.total_ms[i];
            if (count) count[i] = ctx->prof.count[i];
        }
        if (reset)
            for (int i = 0; i < K_COUNT; ++i) {
                ctx->prof.total_ms[i] = 0;
                ctx->prof.count[i] = 0;
            }
        return Q3X_OK;
    } catch (const std::exception& e) {  // nothing crosses the C ABI as an exception
        set_error(std::string("q3x_profile_read: ") + e.what());
        return Q3X_ERR_INTERNAL;
    }
}

const char* q3x_profile_kernel_name(int id) {
    static const char* names[K_COUNT] = {"zsticks_inverse", "plane_sandwich", "zsticks_forward", "kb_overlap_gemm",
                                         "kb_dblock",       "kb_update_gemm", "eig_prepare",     "eig_gemms",
                                         "eig_tridiagonal", "eig_bisection",  "eig_inverse_iteration",
                                         "eig_back_transform"};
    return (id >= 0 && id < K_COUNT) ? names[id] : "";
}

q3x_status q3x_zgemm(q3x_context ctx, char transa, int m, int n, int k, double alpha, const void* A, long lda,
                     const void* B, long ldb, double beta, void* C, long ldc, int precision) {
    try {
        if (!ctx || m < 0 || n < 0 || k < 0 || (transa != 'N' && transa != 'C') ||
            (precision != Q3X_FP64 && precision != Q3X_MIXED)) {
            set_error("q3x_zgemm: invalid argument");
            return Q3X_ERR_ARGUMENT;
        }
        const long rows_a = transa == 'N' ? m : k;
        if (lda < std::max<long>(1, rows_a) || ldb < std::max(1, k) || ldc < std::max(1, m)) {
            set_error("q3x_zgemm: leading dimension too small");
            return Q3X_ERR_ARGUMENT;
        }
        const size_t ws = size_t(64) << 20;
        if (!ctx->gemm_work.ensure(ws)) return Q3X_ERR_DEVICE;
        const GemmTypes t = precision == Q3X_FP64 ? GemmTypes::D_D_D_D : GemmTypes::D_D_F_D;
        return hip_ok(zgemm(t, transa == 'C' ? 1 : 0, m, n, k, alpha, A, lda, B, ldb, beta, C, ldc,
                            ctx->gemm_work.get(), ws, ctx->stream),
                      "q3x_zgemm")
                   ? Q3X_OK
                   : Q3X_ERR_DEVICE;
    } catch (const std::exception& e) {  // nothing crosses the C ABI as an exception
        set_error(std::string("q3x_zgemm: ") + e.what());
        return Q3X_ERR_INTERNAL;
    }
}

q3x_status q3x_hegv_device(q3x_context ctx, int n, const void* H, long ldh, const void* S, long lds, int nvec,
                           double* w, void* X, long ldx) {
    try {
        if (!ctx || n < 0 || nvec < 0 || nvec > n ||
            (n > 0 && (!H || !S || (nvec > 0 && (!w || !X)) || ldh < n || lds < n || ldx < n))) {
            set_error("q3x_hegv_device: invalid argument");
            return Q3X_ERR_ARGUMENT;
        }
        if (n == 0 || nvec == 0) return Q3X_OK;
        const size_t ws = size_t(64) << 20;
        if (!ctx->gemm_work.ensure(ws) || !ctx->eig_work.ensure(subspace_eig_workspace(n, nvec))) return Q3X_ERR_DEVICE;
        hipError_t e = hipSuccess;
        const int info = subspace_eig(n, static_cast<const double*>(H), ldh, static_cast<const double*>(S), lds,
                                      nvec, w,
                                      static_cast<double*>(X), ldx, ctx->eig_work.get(), ctx->eig_work.size(),
                                      ctx->gemm_work.get(), ws, ctx->max_lds, ctx->num_cus, ctx->stream, &e, nullptr,
                                      &ctx->prof, &ctx->pin);
        if (info < 0) return hip_ok(e, "q3x_hegv_device") ? Q3X_ERR_INTERNAL : Q3X_ERR_DEVICE;
        if (info > 0) {
            set_error("q3x_hegv_device: S is not positive definite");
            return Q3X_ERR_NUMERIC;
        }
        return Q3X_OK;
    } catch (const std::exception& e) {  // nothing crosses the C ABI as an exception
        set_error(std::string("q3x_hegv_device: ") + e.what());
        return Q3X_ERR_INTERNAL;
    }
}

void q3x_debug_force_barrier_giveups(int count) { force_barrier_giveups(count); }

void q3x_debug_fail_eigensolves(unsigned long long mask) { fail_eigensolves(mask); }

long q3x_eig_barrier_redos_total(void) { return barrier_redos_total(); }

q3x_status q3x_device_alloc(q3x_context ctx, unsigned long bytes, void** out) {
    try {
        if (!ctx || !out) {
            set_error("q3x_device_alloc: invalid argument");
            return Q3X_ERR_ARGUMENT;
        }
        return hip_ok(hipMalloc(out, bytes ? bytes : 16), "hipMalloc") ? Q3X_OK : Q3X_ERR_DEVICE;
    } catch (const std::exception& e) {  // nothing crosses the C ABI as an exception
        set_error(std::string("q3x_device_alloc: ") + e.what());
        return Q3X_ERR_INTERNAL;
    }
}

q3x_status q3x_device_copy(q3x_context ctx, void* dst, const void* src, unsigned long bytes) {
    try {
        if (!ctx || (bytes % 16) != 0) {
            set_error("q3x_device_copy: bytes must be a multiple of 16");
            return Q3X_ERR_ARGUMENT;
        }
        return hip_ok(launch_copy16(src, dst, static_cast<long>(bytes / 16), ctx->stream), "copy kernel")
                   ? Q3X_OK
                   : Q3X_ERR_DEVICE;
    } catch (const std::exception& e) {  // nothing crosses the C ABI as an exception
        set_error(std::string("q3x_device_copy: ") + e.what());
        return Q3X_ERR_INTERNAL;
    }
}

q3x_status q3x_device_free(q3x_context ctx, void* p) {
    try {
        (void)ctx;
        return hip_ok(hipFree(p), "hipFree") ? Q3X_OK : Q3X_ERR_DEVICE;
    } catch (const std::exception& e) {  // nothing crosses the C ABI as an exception
        set_error(std::string("q3x_device_free: ") + e.what());
        return Q3X_ERR_INTERNAL;
    }
}

q3x_status q3x_copy_to_device(q3x_context ctx, void* dst, const void* src, unsigned long bytes) {
    try {
        if (!ctx) return Q3X_ERR_ARGUMENT;
        return hip_ok(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, ctx->stream), "H2D") &&
                       hip_ok(hipStreamSynchronize(ctx->stream), "sync")
                   ? Q3X_OK
                   : Q3X_ERR_DEVICE;
    } catch (const std::exception& e) {  // nothing crosses the C ABI as an exception
        set_error(std::string("q3x_copy_to_device: ") + e.what());
        return Q3X_ERR_INTERNAL;
    }
}

q3x_status q3x_copy_to_host(q3x_context ctx, void* dst, const void* src, unsigned long bytes) {
    try {
        if (!ctx) return Q3X_ERR_ARGUMENT;
        return hip_ok(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, ctx->stream), "D2H") &&
                       hip_ok(hipStreamSynchronize(ctx->stream), "sync")
                   ? Q3X_OK
                   : Q3X_ERR_DEVICE;
    } catch (const std::exception& e) {  // nothing crosses the C ABI as an exception
        set_error(std::string("q3x_copy_to_host: ") + e.what());
        return Q3X_ERR_INTERNAL;
    }
}


q3x_status q3x_hamiltonian_create(q3x_context ctx, q3x_basis b, q3x_hamiltonian* out) {
    try {
        if (!ctx || !b || !out) {
            set_error("q3x_hamiltonian_create: invalid argument");
            return Q3X_ERR_ARGUMENT;
        }
        auto* h = new q3x_hamiltonian_s;
        h->ctx = ctx;
        const q3x_status s = hamiltonian_rebind(h, b);
        if (s != Q3X_OK) {
            delete h;
            return s;
        }
        *out = h;
        return Q3X_OK;
    } catch (const std::exception& e) {  // nothing crosses the C ABI as an exception
        set_error(std::string("q3x_hamiltonian_create: ") + e.what());
        return Q3X_ERR_INTERNAL;
    }
}

q3x_status q3x_hamiltonian_destroy(q3x_hamiltonian h) {
    try {
        if (h && h->ctx) (void)hipStreamSynchronize(h->ctx->stream);
        delete h;
        return Q3X_OK;
    } catch (const std::exception& e) {  // nothing crosses the C ABI as an exception
        set_error(std::string("q3x_hamiltonian_destroy: ") + e.what());
        return Q3X_ERR_INTERNAL;
    }
}

q3x_status q3x_hamiltonian_set_potential(q3x_hamiltonian h, const double* v, long n) {
    try {
        if (!h || !v) {
            set_error("set_potential: invalid argument");
            return Q3X_ERR_ARGUMENT;
        }
        const int* rd = h->basis->d.rho;
        const long nrho = static_cast<long>(rd[0]) * rd[1] * rd[2];
        const bool same_grid = rd[0] == h->dims[0] && rd[1] == h->dims[1] && rd[2] == h->dims[2];
        if (n != nrho && n != h->npoints) {
            set_error("set_potential: V must hold the density grid's n1*n2*n3 points (or the wavefunction grid's)");
            return Q3X_ERR_ARGUMENT;
        }
        hipStream_t st = h->ctx->stream;
        std::vector<double> restricted;
        const double* vs = v;  // V on the wavefunction grid
        if (n == nrho && !same_grid) {
            // compute_smooth_grid_potential (serial.hpp:852-872): G-space restriction on the device
            FftTables dense[3];
            for (int d = 0; d < 3; ++d)
                if (!build_tables(dense[d], rd[d], st, &h->ctx->pin)) return Q3X_ERR_DEVICE;
            DevBuf vd, wd, ws, vo;
            if (!vd.ensure(static_cast<size_t>(nrho) * 8) || !wd.ensure(static_cast<size_t>(nrho) * 16) ||
                !ws.ensure(static_cast<size_t>(h->npoints) * 16) || !vo.ensure(static_cast<size_t>(h->npoints) * 8))
                return Q3X_ERR_DEVICE;
            if (!hip_ok(hipMemcpyAsync(vd.get(), v, static_cast<size_t>(nrho) * 8, hipMemcpyHostToDevice, st), "H2D V"))
                return Q3X_ERR_DEVICE;
            const FftPlan dp[3] = {dense[0].plan, dense[1].plan, dense[2].plan};
            const FftPlan sp[3] = {h->fft[0].plan, h->fft[1].plan, h->fft[2].plan};
            const void* dtw[3] = {dense[0].tw64.get(), dense[1].tw64.get(), dense[2].tw64.get()};
            const void* stw[3] = {h->fft[0].tw64.get(), h->fft[1].tw64.get(), h->fft[2].tw64.get()};
            if (!hip_ok(launch_restrict_potential(vd.as<double>(), rd, dp, dtw, h->dims, sp, stw, wd.get(), ws.get(),
                                                  vo.as<double>(), st),
                        "potential restriction kernels"))
                return Q3X_ERR_DEVICE;
            restricted.resize(h->npoints);
            if (!hip_ok(hipMemcpyAsync(restricted.data(), vo.get(), static_cast<size_t>(h->npoints) * 8,
                                       hipMemcpyDeviceToHost, st),
                        "D2H smooth V") ||
                !hip_ok(hipStreamSynchronize(st), "hipStreamSynchronize"))
                return Q3X_ERR_DEVICE;
            vs = restricted.data();
        }
        const int n1 = h->dims[0], n2 = h->dims[1], n3 = h->dims[2];
        const double inv = 1.0 / static_cast<double>(h->npoints);  // serial.hpp:69-73
        std::vector<double> d(h->npoints);
        std::vector<float> f(h->npoints);
        for (int i1 = 0; i1 < n1; ++i1)
            for (int i2 = 0; i2 < n2; ++i2)
                for (int i3 = 0; i3 < n3; ++i3) {
                    const double s = vs[(static_cast<long>(i1) * n2 + i2) * n3 + i3] * inv;
                    const long o = (static_cast<long>(i3) * n2 + i2) * n1 + i1;
                    d[o] = s;
                    f[o] = static_cast<float>(s);
                }
        if (!h->v64.upload(d.data(), d.size() * sizeof(double), h->ctx->stream, &h->ctx->pin) ||
            !h->v32.upload(f.data(), f.size() * sizeof(float), h->ctx->stream, &h->ctx->pin))
            return Q3X_ERR_DEVICE;
        h->have_potential = true;
        return Q3X_OK;
    } catch (const std::exception& e) {  // nothing crosses the C ABI as an exception
        set_error(std::string("q3x_hamiltonian_set_potential: ") + e.what());
        return Q3X_ERR_INTERNAL;
    }
}

q3x_status q3x_hamiltonian_prepare_kpoint(q3x_hamiltonian h, q3x_kpoint k) {
    try {
        if (!h || !k || k->basis != h->basis) {
            set_error("prepare_kpoint: k-point must come from the Hamiltonian's basis");
            return Q3X_ERR_ARGUMENT;
        }
        const KPointData& K = k->d;
        const StickLayout& S = K.sticks;
        hipStream_t st = h->ctx->stream;
        h->npw = static_cast<int>(K.gidx.size());
        h->gamma = K.gamma;
        // the list the transforms and the kinetic term run over (Gamma: half sphere + mirrors)
        const std::vector<int>& tmap = K.gamma ? K.ext_fftmap : K.fftmap;
        const std::vector<double>& tkin = K.gamma ? K.ext_ekin : K.ekin;
        h->npw_ext = static_cast<int>(tmap.size());
        h->nst = S.nst;
        h->ncol = S.ncol;
        std::vector<int> out;
        for (int p = 0; p < h->npw_ext; ++p)
            if (tmap[p] < 0) out.push_back(p);
        h->n_outside = static_cast<int>(out.size());
        const int n1 = h->dims[0], n2 = h->dims[1], n3 = h->dims[2];
        h->sticks_per_block = std::max(4, std::min(64, 3072 / n3));
        // two-level plane kernel where the plane is square and of a size it is built for;
        // Q3X_FFT_KERNELS=generic keeps the Stockham kernels (used to compare the two)
        const char* pk = std::getenv("Q3X_FFT_KERNELS");
        const bool use_tl = !(pk && std::strcmp(pk, "generic") == 0) && plane_tl_supported(n1, n2);
        h->tl64 = use_tl ? plane_tl_config(n1, S.ncol, false, h->ctx->max_lds, h->ctx->num_cus) : PlaneTlConfig{};
        h->tl32 = use_tl ? plane_tl_config(n1, S.ncol, true, h->ctx->max_lds, h->ctx->num_cus) : PlaneTlConfig{};
        const bool use_tz = !(pk && std::strcmp(pk, "generic") == 0) && stick_tl_supported(n3);
        h->tz64 = use_tz ? stick_tl_config(n3, false) : StickTlConfig{};
        h->tz32 = use_tz ? stick_tl_config(n3, true) : StickTlConfig{};
        // generic plane kernel LDS: twiddles + [ncol][n2] plane + a buffer of rows_per_pass rows of n1 for each of
        // 8 waves
        for (int fp32 = 0; fp32 < 2; ++fp32) {
            const size_t cs = fp32 ? 8 : 16;
            const int ps = fp32 ? lds_pad_shift<cx<float>>() : lds_pad_shift<cx<double>>();
            const size_t l1 = padded_len(n1, ps), l2 = padded_len(n2, ps);
            const size_t fixed = cs * (n1 + n2 + l2 * S.ncol);
            long rp = 0;
            if (fixed < h->ctx->max_lds) rp = static_cast<long>((h->ctx->max_lds - fixed) / (cs * l1 * 8));
            rp = std::min<long>(rp, 4);
            rp = std::min<long>(rp, n2);
            if (rp < 1 && !(fp32 ? h->tl32.ok : h->tl64.ok)) {
                set_error("prepare_kpoint: the occupied plane does not fit in LDS for this box");
                return Q3X_ERR_ARGUMENT;
            }
            (fp32 ? h->rows_per_pass32 : h->rows_per_pass64) = static_cast<int>(std::max<long>(rp, 1));
        }
        bool ok = h->stick_off.upload(S.stick_off.data(), S.stick_off.size() * sizeof(int), st, &h->ctx->pin) &&
                  h->stick_i2.upload(S.stick_i2.data(), S.stick_i2.size() * sizeof(int), st, &h->ctx->pin) &&
                  h->stick_col.upload(S.stick_col.data(), S.stick_col.size() * sizeof(int), st, &h->ctx->pin) &&
                  h->col_i1.upload(S.col_i1.data(), S.col_i1.size() * sizeof(int), st, &h->ctx->pin) &&
                  h->ent_pw.upload(S.ent_pw.data(), S.ent_pw.size() * sizeof(int), st, &h->ctx->pin) &&
                  h->ent_pos.upload(S.ent_pos.data(), S.ent_pos.size() * sizeof(int), st, &h->ctx->pin) &&
                  h->ekin.upload(tkin.data(), tkin.size() * sizeof(double), st, &h->ctx->pin) &&
                  h->outside.upload(out.data(), out.size() * sizeof(int), st, &h->ctx->pin);
        if (!ok) return Q3X_ERR_DEVICE;
        h->kp = k;
        h->have_projectors = false;
        h->nproj = 0;
        return Q3X_OK;
    } catch (const std::exception& e) {  // nothing crosses the C ABI as an exception
        set_error(std::string("q3x_hamiltonian_prepare_kpoint: ") + e.what());
        return Q3X_ERR_INTERNAL;
    }
}

}  // extern "C"

namespace q3x {
// Per-atom D blocks of a projector set (serial.hpp:587-679): the packed blocks, and per projector
// its block's first projector, size and offset; marks the set complete.
q3x_status upload_dblocks(q3x_hamiltonian h, int nproj, int natoms, const int* block_sizes, const double* d_blocks) {
    hipStream_t st = h->ctx->stream;
    std::vector<int> ps(nproj), pz(nproj);
    std::vector<long> po(nproj);
    std::vector<double> D;
    int start = 0;
    long off = 0;
    for (int a = 0; a < natoms; ++a) {
        const int bs = block_sizes[a];
        for (int q = 0; q < bs * bs; ++q) {
            const double dv = d_blocks[off + q];
            D.push_back(std::abs(dv) > 1e-15 ? dv : 0.0);  // d_matrix_element_threshold
        }
        for (int i = 0; i < bs; ++i) {
            ps[start + i] = start;
            pz[start + i] = bs;
            po[start + i] = off;
        }
        start += bs;
        off += static_cast<long>(bs) * bs;
    }
    if (D.empty()) D.push_back(0.0);
    const bool ok = h->dblocks.upload(D.data(), D.size() * sizeof(double), st, &h->ctx->pin) &&
                    h->pstart.upload(ps.data(), ps.size() * sizeof(int), st, &h->ctx->pin) &&
                    h->psize.upload(pz.data(), pz.size() * sizeof(int), st, &h->ctx->pin) &&
                    h->poff.upload(po.data(), po.size() * sizeof(long), st, &h->ctx->pin);
    if (!ok) return Q3X_ERR_DEVICE;
    h->nproj = nproj;
    h->have_projectors = true;
    return Q3X_OK;
}
}  // namespace q3x

extern "C" {

q3x_status q3x_hamiltonian_set_projectors(q3x_hamiltonian h, const double* beta, int nproj, int natoms,
                                          const int* block_sizes, const double* d_blocks) {
    try {
        if (!h || !h->kp) {
            set_error("set_projectors: prepare_kpoint() first");
            return Q3X_ERR_STATE;
        }
        if (nproj < 0 || natoms < 0 || (nproj > 0 && (!beta || !block_sizes || !d_blocks))) {
            set_error("set_projectors: invalid argument");
            return Q3X_ERR_ARGUMENT;
        }
        int total = 0;
        for (int a = 0; a < natoms; ++a) total += block_sizes[a];
        if (total != nproj) {
            set_error("set_projectors: block sizes must sum to nproj");
            return Q3X_ERR_ARGUMENT;
        }
        hipStream_t st = h->ctx->stream;
        const size_t n = static_cast<size_t>(h->npw) * nproj;
        std::vector<float> b32(2 * n);
        for (size_t i = 0; i < 2 * n; ++i) b32[i] = static_cast<float>(beta[i]);
        bool ok = h->beta64.upload(beta, n * 16, st, &h->ctx->pin) &&
                  h->beta32.upload(b32.data(), n * 8, st, &h->ctx->pin);
        if (!ok) return Q3X_ERR_DEVICE;
        return upload_dblocks(h, nproj, natoms, block_sizes, d_blocks);
    } catch (const std::exception& e) {  // nothing crosses the C ABI as an exception
        set_error(std::string("q3x_hamiltonian_set_projectors: ") + e.what());
        return Q3X_ERR_INTERNAL;
    }
}

q3x_status q3x_hamiltonian_apply_device(q3x_hamiltonian h, const void* d_psi, void* d_hpsi, int nvec,
                                        int precision) {
    try {
        if (!h) {
            set_error("apply_device: null Hamiltonian");
            return Q3X_ERR_ARGUMENT;
        }
        return hamiltonian_apply_device(h, static_cast<const double*>(d_psi), h->npw, static_cast<double*>(d_hpsi),
                                        h->npw,
                                        nvec, precision);
    } catch (const std::exception& e) {  // nothing crosses the C ABI as an exception
        set_error(std::string("q3x_hamiltonian_apply_device: ") + e.what());
        return Q3X_ERR_INTERNAL;
    }
}

q3x_status q3x_hamiltonian_accumulate_density(q3x_hamiltonian h, const void* d_psi, int nb, const double* weights,
                                              void* d_rho) {
    try {
        if (!h || !d_psi || !weights || !d_rho || nb < 0) {
            set_error("q3x_hamiltonian_accumulate_density: invalid argument");
            return Q3X_ERR_ARGUMENT;
        }
        if (!h->kp) {
            set_error("q3x_hamiltonian_accumulate_density: prepare_kpoint() must come first");
            return Q3X_ERR_STATE;
        }
        if (!h->tl64.ok || h->tl64.plane_global || !h->tz64.ok || h->n_outside > 0) {
            set_error("q3x_hamiltonian_accumulate_density: the wavefunction box must be one the two-level kernels "
                      "cover (square planes of a listed size, the plane in LDS) and hold every plane wave");
            return Q3X_ERR_ARGUMENT;
        }
        if (nb == 0) return Q3X_OK;
        hipStream_t st = h->ctx->stream;
        const HLocalArgs args = make_args(h, false);
        const int ncols = h->gamma ? (nb + 1) / 2 : nb;
        // per column (w_re, w_im): a band's weight twice, or the weights of a packed Gamma pair
        std::vector<double> w2(2 * static_cast<size_t>(ncols));
        for (int c = 0; c < ncols; ++c) {
            if (h->gamma) {
                w2[2 * c] = weights[2 * c];
                w2[2 * c + 1] = 2 * c + 1 < nb ? weights[2 * c + 1] : 0.0;
            } else {
                w2[2 * c] = w2[2 * c + 1] = weights[c];
            }
        }
        if (!h->dens_w.upload(w2.data(), w2.size() * sizeof(double), st, &h->ctx->pin)) return Q3X_ERR_DEVICE;
        const int chunk = band_chunk(h, h->tl64, ncols);
        if (!h->zbuf.ensure(static_cast<size_t>(chunk) * h->dims[2] * h->nst * 16)) return Q3X_ERR_DEVICE;
        if (h->gamma && !h->gpack.ensure(static_cast<size_t>(h->npw_ext) * chunk * 16)) return Q3X_ERR_DEVICE;
        const double* psi = static_cast<const double*>(d_psi);
        for (int c0 = 0; c0 < ncols; c0 += chunk) {
            const int nc = std::min(chunk, ncols - c0);
            const double* in = psi + 2L * c0 * h->npw;
            long ldin = h->npw;
            if (h->gamma) {  // bands 2 c0 .. packed in pairs on the extended list (accumulator.hpp:52-56)
                const int nbands = std::min(2 * nc, nb - 2 * c0);
                if (!hip_ok(launch_gamma_pack(h->npw, nbands, psi + 4L * c0 * h->npw, h->npw, h->gpack.as<double>(),
                                              h->npw_ext, st),
                            "Gamma pack"))
                    return Q3X_ERR_DEVICE;
                in = h->gpack.as<double>();
                ldin = h->npw_ext;
            }
            if (!hip_ok(launch_sticks_tl<double>(args, +1, in, ldin, h->zbuf.as<cx<double>>(), nullptr, 0, h->tz64, nc,
                                                 st),
                        "density z-sticks") ||
                !hip_ok(launch_plane_density(args, h->zbuf.as<cx<double>>(), h->tl64, nc, h->dens_w.as<double>() + 2L * c0,
                                             static_cast<double*>(d_rho), st),
                        "density planes"))
                return Q3X_ERR_DEVICE;
        }
        return Q3X_OK;
    } catch (const std::exception& e) {  // nothing crosses the C ABI as an exception
        set_error(std::string("q3x_hamiltonian_accumulate_density: ") + e.what());
        return Q3X_ERR_INTERNAL;
    }
}

q3x_status q3x_hamiltonian_apply(q3x_hamiltonian h, const double* psi, double* hpsi, int nvec, int precision) {
    try {
        if (!h || !h->kp) {
            set_error("apply: prepare_kpoint() must be called before apply (serial.hpp:487-492)");
            return Q3X_ERR_STATE;
        }
        if (nvec < 0 || (nvec > 0 && (!psi || !hpsi))) {
            set_error("apply: invalid argument");
            return Q3X_ERR_ARGUMENT;
        }
        if (nvec == 0) return Q3X_OK;
        hipStream_t st = h->ctx->stream;
        const size_t bytes = static_cast<size_t>(h->npw) * nvec * 16;
        if (!h->io_psi.ensure(bytes) || !h->io_hpsi.ensure(bytes)) return Q3X_ERR_DEVICE;
        if (!hip_ok(hipMemcpyAsync(h->io_psi.get(), psi, bytes, hipMemcpyHostToDevice, st), "H2D psi"))
            return Q3X_ERR_DEVICE;
        q3x_status s = hamiltonian_apply_device(h, h->io_psi.as<double>(), h->npw, h->io_hpsi.as<double>(), h->npw,
                                                nvec,
                                                precision);
        if (s != Q3X_OK) return s;
        if (!hip_ok(hipMemcpyAsync(hpsi, h->io_hpsi.get(), bytes, hipMemcpyDeviceToHost, st), "D2H hpsi") ||
            !hip_ok(hipStreamSynchronize(st), "hipStreamSynchronize"))
            return Q3X_ERR_DEVICE;
        return Q3X_OK;
    } catch (const std::exception& e) {  // nothing crosses the C ABI as an exception
        set_error(std::string("q3x_hamiltonian_apply: ") + e.what());
        return Q3X_ERR_INTERNAL;
    }
}

}  // extern "C"
