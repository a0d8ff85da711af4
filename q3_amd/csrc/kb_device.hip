// Kleinman-Bylander projector set of the prepared k-point built on the device, straight into the
// applicator's projector buffers (the product of GPUProjectorSetBuilder,
// hamiltonian_applicator/projector_set_builder_gpu.hpp:89-300, with every factor formed on the
// device too: only the beta(q) tables, the G list of the k-point and the atoms go up).
//
//   beta_p(k+G) = Omega^-1/2 beta_l(|k+G|) Y_lm(k+G) exp(-i 2 pi (k+G).tau) (-i)^l
//
// Kernel 1 (one thread per plane wave): q = |k+G|, the real Y_lm up to lmax and the form factor
// beta_l(q) Y_lm(k+G) of every channel of every species (4-point Lagrange on the table), the
// same arithmetic in the same order as the host builder.  Kernel 2 (plane wave x atom): the
// atom's structure factor times its species' form factors times Omega^-1/2 (-i)^l, written as
// complex128 and complex64 columns (atom, beta, m = 0, +1, -1, ...).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <string>
#include <vector>

#include "kb_plan.h"
#include "runtime.h"

namespace q3x {
namespace {

constexpr double kPi = 3.14159265358979323846;
constexpr double kTwoPi = 2.0 * kPi;
constexpr double kFourPi = 4.0 * kPi;
constexpr double kDq = 0.01;  // default_q_grid_spacing (tolerances.hpp:109)

__device__ double interp4_dev(const double* t, int nq, double q) {
    if (q < 1e-10) return t[0];
    const double fi = q / kDq;
    const int b = static_cast<int>(fi);
    const double u = fi - b;
    if (b + 3 >= nq) return 0.0;
    const double w1 = 1.0 - u, w2 = 2.0 - u, w3 = 3.0 - u;
    return t[b] * w1 * w2 * w3 / 6.0 + t[b + 1] * u * w2 * w3 / 2.0 - t[b + 2] * u * w1 * w3 / 2.0 +
           t[b + 3] * u * w1 * w2 / 6.0;
}

// Real Y_lm, l <= 3 (16 slots), by the host builder's normalised associated-Legendre recursion
// (harmonics.hpp:217-327).
__device__ void real_ylm_dev(int lmax, const double v[3], double v2, double* out) {
    const double vm = sqrt(v2);
    const bool flat = vm < 1e-9;
    const double ct = flat ? 0.0 : v[2] / vm;
    const double st = sqrt(fmax(0.0, 1.0 - ct * ct));
    const double phi = flat ? copysign(kPi / 2.0, v[1]) : atan2(v[1], v[0]);
    double Q[10];  // (l, m) at l (l + 1) / 2 + m, l <= 3
    for (int i = 0; i < 10; ++i) Q[i] = 0.0;
    Q[0] = 1.0;
    if (lmax >= 1) {
        Q[1] = ct;
        Q[2] = -st / sqrt(2.0);
    }
    for (int l = 2; l <= lmax; ++l) {
        for (int m = 0; m <= l - 2; ++m) {
            const double den = sqrt(static_cast<double>(l * l - m * m));
            Q[l * (l + 1) / 2 + m] = ct * (2 * l - 1) / den * Q[(l - 1) * l / 2 + m] -
                                     sqrt(static_cast<double>((l - 1) * (l - 1) - m * m)) / den *
                                         Q[(l - 2) * (l - 1) / 2 + m];
        }
        const double df = sqrt(static_cast<double>(2 * l - 1));
        const double prev = Q[(l - 1) * l / 2 + l - 1];
        Q[l * (l + 1) / 2 + l - 1] = ct * df * prev;
        Q[l * (l + 1) / 2 + l] = -df / sqrt(static_cast<double>(2 * l)) * st * prev;
    }
    out[0] = Q[0] / sqrt(kFourPi);
    for (int l = 1; l <= lmax; ++l) {
        const double nrm = sqrt(static_cast<double>(2 * l + 1) / kFourPi);
        out[l * l] = nrm * Q[l * (l + 1) / 2];
        const double off = nrm * sqrt(2.0);
        for (int m = 1; m <= l; ++m) {
            out[l * l + 2 * m - 1] = off * Q[l * (l + 1) / 2 + m] * cos(m * phi);
            out[l * l + 2 * m] = off * Q[l * (l + 1) / 2 + m] * sin(m * phi);
        }
    }
}

// F[c][p] = beta_{l(c)}(|k+G_p|) Y_{slot(c)}(k+G_p) for the nchan channels of all species.
__global__ void kb_form_factors(int npw, const double* __restrict__ g, double kc0, double kc1, double kc2, double tpa,
                                int lmax, int nchan, const int* __restrict__ chan_table,
                                const int* __restrict__ chan_slot, const double* __restrict__ tables, int nq,
                                double* __restrict__ F) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= npw) return;
    const double v[3] = {kc0 + g[3 * p], kc1 + g[3 * p + 1], kc2 + g[3 * p + 2]};
    const double v2 = v[0] * v[0] + v[1] * v[1] + v[2] * v[2];
    const double q = sqrt(v2) * tpa;
    double ylm[16];
    real_ylm_dev(lmax, v, v2, ylm);
    for (int c = 0; c < nchan; ++c) {
        const double radial = interp4_dev(tables + static_cast<long>(chan_table[c]) * nq, nq, q);
        F[static_cast<long>(c) * npw + p] = radial * ylm[chan_slot[c]];
    }
}

// beta[:, col0(a) + c] = F[first(s) + c] exp(-i 2 pi ((g . tau_a) + k . tau_a)) Omega^-1/2 (-i)^l.
__global__ void kb_products(int npw, const double* __restrict__ g, const double* __restrict__ tau,
                            const double* __restrict__ kdt, const int* __restrict__ atom_species,
                            const int* __restrict__ atom_col0, const int* __restrict__ spec_first,
                            const int* __restrict__ spec_nchan, const int* __restrict__ chan_l,
                            const double* __restrict__ F, double inv_sqrt_vol, double* __restrict__ beta64,
                            float* __restrict__ beta32) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    const int a = blockIdx.y;
    if (p >= npw) return;
    const double* t = tau + 3 * a;
    const double ph = kTwoPi * ((g[3 * p] * t[0] + g[3 * p + 1] * t[1] + g[3 * p + 2] * t[2]) + kdt[a]);
    double sn, cs;
    sincos(ph, &sn, &cs);
    const double pr = cs, pi = -sn;  // exp(-i ph)
    const int s = atom_species[a];
    const int c0 = spec_first[s], nc = spec_nchan[s];
    const long col0 = atom_col0[a];
    for (int c = 0; c < nc; ++c) {
        const double f = F[static_cast<long>(c0 + c) * npw + p];
        const double zr = f * pr, zi = f * pi;
        // (-i)^l Omega^-1/2: one of its parts is zero
        const int l4 = chan_l[c0 + c] & 3;
        const double ar = l4 == 0 ? inv_sqrt_vol : (l4 == 2 ? -inv_sqrt_vol : 0.0);
        const double ai = l4 == 1 ? -inv_sqrt_vol : (l4 == 3 ? inv_sqrt_vol : 0.0);
        const double br = zr * ar - zi * ai, bi = zr * ai + zi * ar;
        const long o = 2 * ((col0 + c) * npw + p);
        beta64[o] = br;
        beta64[o + 1] = bi;
        beta32[o] = static_cast<float>(br);
        beta32[o + 1] = static_cast<float>(bi);
    }
}

inline unsigned blocks(long n, int t) { return static_cast<unsigned>((n + t - 1) / t); }

}  // namespace
}  // namespace q3x

using namespace q3x;

extern "C" q3x_status q3x_hamiltonian_build_projectors(q3x_hamiltonian h, int natoms, const double* atom_pos,
                                                       const int* atom_species, int nspecies, const int* nmesh,
                                                       const double* r_all, const double* rab_all, const int* nbeta,
                                                       const int* l_all, const int* cutoff_all,
                                                       const double* beta_r_all, const int* block_sizes,
                                                       const double* d_blocks) {
    try {
        if (!h || !h->kp) {
            set_error("build_projectors: prepare_kpoint() first");
            return Q3X_ERR_STATE;
        }
        if (natoms > 0 && (!block_sizes || !d_blocks)) {
            set_error("build_projectors: invalid argument");
            return Q3X_ERR_ARGUMENT;
        }
        const BasisData& B = h->kp->basis->d;
        const KPointData& K = h->kp->d;
        KbPlan plan;
        q3x_status s = kb_plan(B, K, natoms, atom_pos, atom_species, nspecies, nmesh, r_all, rab_all, nbeta, l_all,
                               cutoff_all, beta_r_all, plan);
        if (s != Q3X_OK) return s;
        int total = 0;
        for (int a = 0; a < natoms; ++a) total += block_sizes[a];
        if (total != plan.nproj) {
            set_error("build_projectors: block sizes must sum to the projector count of the species");
            return Q3X_ERR_ARGUMENT;
        }
        const int npw = h->npw, nproj = plan.nproj;
        hipStream_t st = h->ctx->stream;
        if (nproj > 0 && npw > 0) {
            // channels: species-major, beta, m; the tables flattened [species][beta][q]
            std::vector<int> chan_table, chan_slot, chan_l, spec_first(nspecies), spec_nchan(nspecies);
            std::vector<double> tables;
            int tix = 0;
            for (int sp = 0; sp < nspecies; ++sp) {
                spec_first[sp] = static_cast<int>(chan_table.size());
                for (size_t i = 0; i < plan.beta_l[sp].size(); ++i, ++tix) {
                    const int l = plan.beta_l[sp][i];
                    const std::vector<double>& t = (*plan.tables)[sp][i];
                    tables.insert(tables.end(), t.begin(), t.end());
                    for (int mi = 0; mi < 2 * l + 1; ++mi) {
                        chan_table.push_back(tix);
                        chan_slot.push_back(kb_lm_slot(l, kb_channel_m(mi)));
                        chan_l.push_back(l);
                    }
                }
                spec_nchan[sp] = static_cast<int>(chan_table.size()) - spec_first[sp];
            }
            const int nchan = static_cast<int>(chan_table.size());
            std::vector<double> g(3 * static_cast<size_t>(npw));
            for (int p = 0; p < npw; ++p)
                for (int d = 0; d < 3; ++d) g[3 * static_cast<size_t>(p) + d] = B.g[K.gidx[p]].g[d];
            // one upload of the small arrays, then g and the tables
            std::vector<int> ints;
            auto put = [&ints](const std::vector<int>& v) {
                const size_t o = ints.size();
                ints.insert(ints.end(), v.begin(), v.end());
                return o;
            };
            const size_t o_ct = put(chan_table), o_cs = put(chan_slot), o_cl = put(chan_l), o_sf = put(spec_first),
                         o_sn = put(spec_nchan), o_as = put(plan.species_of_atom), o_ac = put(plan.col0_of_atom);
            std::vector<double> dbl(plan.tau);
            const size_t o_kdt = dbl.size();
            dbl.insert(dbl.end(), plan.kdt.begin(), plan.kdt.end());
            const size_t n = static_cast<size_t>(npw) * nproj;
            DevBuf& gd = h->kb_g;
            DevBuf& td = h->kb_tables;
            DevBuf& Fd = h->kb_form;
            DevBuf& id = h->kb_ints;
            DevBuf& dd = h->kb_dbl;
            if (!gd.upload(g.data(), g.size() * sizeof(double), st, &h->ctx->pin) ||
                !td.upload(tables.data(), tables.size() * sizeof(double), st, &h->ctx->pin) ||
                !id.upload(ints.data(), ints.size() * sizeof(int), st, &h->ctx->pin) ||
                !dd.upload(dbl.data(), std::max<size_t>(dbl.size(), 1) * sizeof(double), st, &h->ctx->pin) ||
                !Fd.ensure(static_cast<size_t>(nchan) * npw * sizeof(double)) || !h->beta64.ensure(n * 16) ||
                !h->beta32.ensure(n * 8))
                return Q3X_ERR_DEVICE;
            const int* I = id.as<int>();
            const double* Dv = dd.as<double>();
            const Lattice& L = B.lat;
            hipLaunchKernelGGL(kb_form_factors, dim3(blocks(npw, 256)), dim3(256), 0, st, npw, gd.as<double>(),
                               plan.kc[0], plan.kc[1], plan.kc[2], L.tpa, plan.lmax, nchan, I + o_ct, I + o_cs,
                               td.as<double>(), plan.nq, Fd.as<double>());
            if (!hip_ok(hipGetLastError(), "KB form factors")) return Q3X_ERR_DEVICE;
            if (natoms > 0) {
                hipLaunchKernelGGL(kb_products, dim3(blocks(npw, 256), natoms), dim3(256), 0, st, npw,
                                   gd.as<double>(), Dv, Dv + o_kdt, I + o_as, I + o_ac, I + o_sf, I + o_sn, I + o_cl,
                                   Fd.as<double>(), 1.0 / std::sqrt(L.volume), h->beta64.as<double>(),
                                   h->beta32.as<float>());
                if (!hip_ok(hipGetLastError(), "KB products")) return Q3X_ERR_DEVICE;
            }
        }
        return upload_dblocks(h, nproj, natoms, block_sizes, d_blocks);
    } catch (const std::exception& e) {  // nothing crosses the C ABI as an exception
        set_error(std::string("q3x_hamiltonian_build_projectors: ") + e.what());
        return Q3X_ERR_INTERNAL;
    }
}

namespace q3x {
namespace {

__device__ __forceinline__ unsigned long long mix64(unsigned long long x) {
    x += 0x9E3779B97F4A7C15ull;
    x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
    x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
    return x ^ (x >> 31);
}

// The starting block of q3x_kpoint_random_psi, entry for entry (integer hash, exact conversions,
// one correctly rounded division and product: the same bits as the host generator).
__global__ void random_psi_kernel(int npw, int nb, unsigned long long s0, const double* __restrict__ ekin, int gamma,
                                  double* __restrict__ psi) {
    const long idx = static_cast<long>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (idx >= static_cast<long>(npw) * nb) return;
    const int b = static_cast<int>(idx / npw);
    const int g = static_cast<int>(idx - static_cast<long>(b) * npw);
    const double damp = 1.0 / (1.0 + ekin[g]);
    const unsigned long long o = 2ull * (static_cast<unsigned long long>(b) * npw + g);
    const double scale = 1.0 / 4503599627370496.0;  // 2^-52
    const double re = static_cast<double>(mix64(s0 + o) >> 11) * scale - 1.0;
    const double im = static_cast<double>(mix64(s0 + o + 1) >> 11) * scale - 1.0;
    psi[o] = re * damp;
    psi[o + 1] = gamma && g == 0 ? 0.0 : im * damp;
}

unsigned long long mix64_host(unsigned long long x) {
    x += 0x9E3779B97F4A7C15ull;
    x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
    x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
    return x ^ (x >> 31);
}

}  // namespace
}  // namespace q3x

extern "C" q3x_status q3x_hamiltonian_random_psi(q3x_hamiltonian h, int nb, unsigned long long seed, void* d_psi) {
    try {
        if (!h || !h->kp) {
            set_error("random_psi: prepare_kpoint() first");
            return Q3X_ERR_STATE;
        }
        if (nb < 0 || (nb > 0 && !d_psi)) {
            set_error("random_psi: invalid argument");
            return Q3X_ERR_ARGUMENT;
        }
        const long n = static_cast<long>(h->npw) * nb;
        if (n == 0) return Q3X_OK;
        hipLaunchKernelGGL(random_psi_kernel, dim3(blocks(n, 256)), dim3(256), 0, h->ctx->stream, h->npw, nb,
                           mix64_host(seed), h->ekin.as<double>(), h->gamma ? 1 : 0, static_cast<double*>(d_psi));
        return hip_ok(hipGetLastError(), "random psi") ? Q3X_OK : Q3X_ERR_DEVICE;
    } catch (const std::exception& e) {  // nothing crosses the C ABI as an exception
        set_error(std::string("q3x_hamiltonian_random_psi: ") + e.what());
        return Q3X_ERR_INTERNAL;
    }
}
