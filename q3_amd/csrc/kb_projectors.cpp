// Kleinman-Bylander projector set <k+G|beta> at one k-point, built on the host.
//
// beta_p(k+G) = Omega^-1/2 * beta_l(|k+G|) * Y_lm(k+G) * exp(-i 2pi (k+G).tau) * (-i)^l
// (projector.hpp:42-199).  beta_l(q) is interpolated (4-point Lagrange, lagrange.hpp:16-30)
// from a table on q = 0, 0.01, ... up to 1.1 sqrt(ecutrho) (+4 points)
// (beta_function_table/init.hpp:45-52, uniform_q_grid.hpp:23-28), each entry the spherical
// Bessel transform 4 pi \int r beta_r(r) r j_l(qr) dr by Simpson's rule on the radial mesh
// (spherical_bessel.hpp:33-53, simpson.hpp:22-48).  Projector order: atom, beta, then
// m = 0, +1, -1, +2, -2, ... (harmonics.hpp:40-69).
#include <algorithm>
#include <cmath>
#include <condition_variable>
#include <list>
#include <memory>
#include <mutex>
#include <stdexcept>
#include <thread>
#include <vector>

#include "kb_plan.h"
#include "q3x_internal.h"

namespace q3x {
namespace {

constexpr double kPi = 3.14159265358979323846;
constexpr double kTwoPi = 2.0 * kPi;
constexpr double kFourPi = 4.0 * kPi;
constexpr double kDq = 0.01;  // default_q_grid_spacing (tolerances.hpp:109)

// j_l(x): ascending series below x = l + 1, closed forms / upward recurrence above
// (bessel.hpp:29-190).
double jl_series(int l, double x) {
    double xl = l == 0 ? 1.0 : std::pow(x, l);
    double dfact = 1.0;
    for (int i = 2 * l + 1; i >= 1; i -= 2) dfact *= i;
    double h = 0.5 * x * x, term = 1.0, sum = 1.0;
    for (int i = 1; i <= 60; ++i) {
        term *= -h / (i * (2.0 * l + 2.0 * i + 1.0));
        sum += term;
        if (std::abs(term) < 1e-18 * std::abs(sum)) break;
    }
    return (xl / dfact) * sum;
}

double jl(int l, double x) {
    if (std::abs(x) < 1e-8) return l == 0 ? 1.0 : 0.0;
    const double cross = l <= 2 ? 1.0 : l + 1.0;
    if (std::abs(x) < cross) return jl_series(l, x);
    const double s = std::sin(x), c = std::cos(x);
    if (l == 0) return s / x;
    if (l == 1) return (s / x - c) / x;
    if (l == 2) return ((3.0 / (x * x) - 1.0) * s - 3.0 * c / x) / x;
    double jm = s / x, jc = (s / x - c) / x;
    for (int i = 2; i <= l; ++i) {
        double jp = ((2 * i - 1) * jc / x) - jm;
        jm = jc;
        jc = jp;
    }
    return jc;
}

double simpson_rab(const std::vector<double>& f, const double* rab, int n) {
    if (n < 3) return 0.0;
    double s = 0.0;
    for (int i = 1; i < n - 1; ++i) s += ((i & 1) ? 4.0 : 2.0) * f[i] * rab[i];
    if (n & 1)
        s += f[0] * rab[0] + f[n - 1] * rab[n - 1];
    else
        s += f[0] * rab[0] - f[n - 3] * rab[n - 3] * 0.25 + f[n - 2] * rab[n - 2] + f[n - 1] * rab[n - 1] * 1.25;
    return s / 3.0;
}

// Entries are independent (the same arithmetic whatever the thread), so the q points are spread
// over a few threads joined at the end: the table is built once per cell, on the path of its
// first k-point.
std::vector<double> radial_table(const double* r, const double* rab, int cutoff, int l, const double* beta_r,
                                 int nq) {
    std::vector<double> t(nq);
    auto part = [&](int q0, int q1) {
        std::vector<double> f(cutoff);
        for (int iq = q0; iq < q1; ++iq) {
            const double q = iq * kDq;
            for (int i = 0; i < cutoff; ++i) {
                const double j = (std::abs(q) < 1e-8) ? (l == 0 ? 1.0 : 0.0) : jl(l, q * r[i]);
                f[i] = beta_r[i] * r[i] * j;
            }
            t[iq] = kFourPi * simpson_rab(f, rab, cutoff);
        }
    };
    const int nt = std::max(1, std::min({static_cast<int>(std::thread::hardware_concurrency()), 8, nq / 64}));
    std::vector<std::thread> th;
    for (int i = 1; i < nt; ++i) th.emplace_back(part, static_cast<int>(static_cast<long>(nq) * i / nt),
                                                 static_cast<int>(static_cast<long>(nq) * (i + 1) / nt));
    part(0, static_cast<int>(nq / nt));
    for (auto& x : th) x.join();
    return t;
}

double interp4(const std::vector<double>& t, double q) {
    const int nq = static_cast<int>(t.size());
    if (q < 1e-10) return t[0];
    const double fi = q / kDq;
    const int b = static_cast<int>(fi);
    const double u = fi - b;
    if (b + 3 >= nq) return 0.0;
    const double w1 = 1.0 - u, w2 = 2.0 - u, w3 = 3.0 - u;
    return t[b] * w1 * w2 * w3 / 6.0 + t[b + 1] * u * w2 * w3 / 2.0 - t[b + 2] * u * w1 * w3 / 2.0 +
           t[b + 3] * u * w1 * w2 / 6.0;
}

inline int lm_slot(int l, int m) { return m == 0 ? l * l : (m > 0 ? l * l + 2 * m - 1 : l * l - 2 * m); }

// Real Y_lm with the Condon-Shortley phase via normalised associated-Legendre recursion
// (harmonics.hpp:217-327).
void real_ylm(int lmax, const double v[3], double v2, double* out) {
    const double vm = std::sqrt(v2);
    const bool flat = vm < 1e-9;
    const double ct = flat ? 0.0 : v[2] / vm;
    const double st = std::sqrt(std::max(0.0, 1.0 - ct * ct));
    const double phi = flat ? std::copysign(kPi / 2.0, v[1]) : std::atan2(v[1], v[0]);
    std::vector<double> Q((lmax + 1) * (lmax + 2) / 2, 0.0);
    auto at = [](int l, int m) { return l * (l + 1) / 2 + m; };
    Q[at(0, 0)] = 1.0;
    if (lmax >= 1) {
        Q[at(1, 0)] = ct;
        Q[at(1, 1)] = -st / std::sqrt(2.0);
    }
    for (int l = 2; l <= lmax; ++l) {
        for (int m = 0; m <= l - 2; ++m) {
            const double den = std::sqrt(static_cast<double>(l * l - m * m));
            Q[at(l, m)] = ct * (2 * l - 1) / den * Q[at(l - 1, m)] -
                          std::sqrt(static_cast<double>((l - 1) * (l - 1) - m * m)) / den * Q[at(l - 2, m)];
        }
        const double df = std::sqrt(static_cast<double>(2 * l - 1));
        const double prev = Q[at(l - 1, l - 1)];
        Q[at(l, l - 1)] = ct * df * prev;
        Q[at(l, l)] = -df / std::sqrt(static_cast<double>(2 * l)) * st * prev;
    }
    out[0] = Q[at(0, 0)] / std::sqrt(kFourPi);
    for (int l = 1; l <= lmax; ++l) {
        const double nrm = std::sqrt(static_cast<double>(2 * l + 1) / kFourPi);
        out[lm_slot(l, 0)] = nrm * Q[at(l, 0)];
        const double off = nrm * std::sqrt(2.0);
        for (int m = 1; m <= l; ++m) {
            out[lm_slot(l, m)] = off * Q[at(l, m)] * std::cos(m * phi);
            out[lm_slot(l, -m)] = off * Q[at(l, m)] * std::sin(m * phi);
        }
    }
}

}  // namespace
}  // namespace q3x

using namespace q3x;

namespace q3x {
namespace {

using Tables = std::vector<std::vector<std::vector<double>>>;  // [species][beta][q]

// Last few table sets with the inputs they came from (compared in full on lookup); shared by
// all threads (server and batch workers build projectors concurrently).  A set being built is
// listed in `building`, and a thread that needs the same set waits for it instead of building
// it again (the k-points of one cell reach several threads at once).
struct TableCache {
    std::mutex mu;
    std::condition_variable cv;
    std::list<std::pair<std::vector<double>, std::shared_ptr<const Tables>>> entries;
    std::list<std::vector<double>> building;
};
TableCache& table_cache() {
    static TableCache c;
    return c;
}

template <typename Build>
std::shared_ptr<const Tables> tables_for(const std::vector<double>& key, Build build) {
    TableCache& c = table_cache();
    std::unique_lock<std::mutex> g(c.mu);
    for (;;) {
        for (auto it = c.entries.begin(); it != c.entries.end(); ++it)
            if (it->first == key) {
                c.entries.splice(c.entries.begin(), c.entries, it);  // most recently used first
                return c.entries.front().second;
            }
        if (std::find(c.building.begin(), c.building.end(), key) == c.building.end()) break;
        c.cv.wait(g);
    }
    c.building.push_back(key);
    g.unlock();
    std::shared_ptr<const Tables> t;
    try {
        t = build();
    } catch (...) {
        g.lock();
        c.building.remove(key);
        c.cv.notify_all();
        throw;
    }
    g.lock();
    c.building.remove(key);
    c.entries.emplace_front(key, t);
    while (c.entries.size() > 8) c.entries.pop_back();
    c.cv.notify_all();
    return t;
}

}  // namespace

q3x_status kb_plan(const BasisData& B, const KPointData& K, int natoms, const double* atom_pos,
                   const int* atom_species, int nspecies, const int* nmesh, const double* r_all,
                   const double* rab_all, const int* nbeta, const int* l_all, const int* cutoff_all,
                   const double* beta_r_all, KbPlan& plan) {
    if (natoms < 0 || nspecies < 1 || !nmesh || !nbeta || (natoms > 0 && (!atom_pos || !atom_species))) {
        set_error("kb projectors: invalid argument");
        return Q3X_ERR_ARGUMENT;
    }
    const Lattice& L = B.lat;
    struct Sp {
        const double *r, *rab, *beta_r;
        int nmesh, nbeta;
        const int *l, *cut;
    };
    std::vector<Sp> sp(nspecies);
    size_t ro = 0, bo = 0, bro = 0;
    plan.lmax = 0;
    plan.beta_l.assign(nspecies, {});
    for (int s = 0; s < nspecies; ++s) {
        sp[s] = {r_all + ro, rab_all + ro, beta_r_all + bro, nmesh[s], nbeta[s], l_all + bo, cutoff_all + bo};
        if (nbeta[s] < 0 || nmesh[s] < 0) {
            set_error("kb projectors: negative projector or mesh count");
            return Q3X_ERR_ARGUMENT;
        }
        for (int i = 0; i < nbeta[s]; ++i) {
            if (l_all[bo + i] < 0 || l_all[bo + i] > 3 || cutoff_all[bo + i] > nmesh[s]) {
                set_error("kb projectors: projector l must be 0..3 and cutoff <= nmesh");
                return Q3X_ERR_ARGUMENT;
            }
            plan.lmax = std::max(plan.lmax, l_all[bo + i]);
            plan.beta_l[s].push_back(l_all[bo + i]);
        }
        ro += nmesh[s];
        bo += nbeta[s];
        bro += static_cast<size_t>(nbeta[s]) * nmesh[s];
    }
    plan.species_of_atom.assign(atom_species, atom_species + natoms);
    plan.col0_of_atom.resize(natoms);
    plan.nproj = 0;
    for (int a = 0; a < natoms; ++a) {
        const int s = atom_species[a];
        if (s < 0 || s >= nspecies) {
            set_error("kb projectors: atom species out of range");
            return Q3X_ERR_ARGUMENT;
        }
        plan.col0_of_atom[a] = plan.nproj;
        for (int i = 0; i < sp[s].nbeta; ++i) plan.nproj += 2 * sp[s].l[i] + 1;
    }
    plan.nq = static_cast<int>(std::sqrt(B.ecutrho) * 1.1 / kDq) + 4;
    // The beta(q) tables depend on the species data and nq only, not on k: every k-point of a
    // cell (a band path, a batch) reuses them from a small cache keyed by the exact inputs.
    std::vector<double> key;
    key.push_back(plan.nq);
    key.push_back(nspecies);
    for (int s = 0; s < nspecies; ++s) {
        key.push_back(sp[s].nmesh);
        key.push_back(sp[s].nbeta);
        for (int i = 0; i < sp[s].nbeta; ++i) {
            key.push_back(sp[s].l[i]);
            key.push_back(sp[s].cut[i]);
        }
        key.insert(key.end(), sp[s].r, sp[s].r + sp[s].nmesh);
        key.insert(key.end(), sp[s].rab, sp[s].rab + sp[s].nmesh);
        key.insert(key.end(), sp[s].beta_r, sp[s].beta_r + static_cast<size_t>(sp[s].nbeta) * sp[s].nmesh);
    }
    const int nq = plan.nq;
    plan.tables = tables_for(key, [&] {
        auto t = std::make_shared<KbTables>(nspecies);
        for (int s = 0; s < nspecies; ++s)
            for (int i = 0; i < sp[s].nbeta; ++i)
                (*t)[s].push_back(radial_table(sp[s].r, sp[s].rab, sp[s].cut[i], sp[s].l[i],
                                               sp[s].beta_r + static_cast<size_t>(i) * sp[s].nmesh, nq));
        return std::shared_ptr<const KbTables>(t);
    });
    for (int d = 0; d < 3; ++d) plan.kc[d] = K.k[0] * L.b[0][d] + K.k[1] * L.b[1][d] + K.k[2] * L.b[2][d];
    plan.tau.resize(3 * static_cast<size_t>(natoms));
    plan.kdt.resize(natoms);
    for (int a = 0; a < natoms; ++a) {
        double* tau = &plan.tau[3 * static_cast<size_t>(a)];
        for (int d = 0; d < 3; ++d)
            tau[d] = atom_pos[3 * a] * L.a[0][d] + atom_pos[3 * a + 1] * L.a[1][d] + atom_pos[3 * a + 2] * L.a[2][d];
        plan.kdt[a] = plan.kc[0] * tau[0] + plan.kc[1] * tau[1] + plan.kc[2] * tau[2];
    }
    return Q3X_OK;
}

}  // namespace q3x

extern "C" q3x_status q3x_kb_projectors(q3x_basis b, q3x_kpoint k, int natoms, const double* atom_pos,
                                        const int* atom_species, int nspecies, const int* nmesh,
                                        const double* r_all, const double* rab_all, const int* nbeta,
                                        const int* l_all, const int* cutoff_all, const double* beta_r_all,
                                        double* beta_out, int* nproj_out) {
    try {
        if (!b || !k) {
            set_error("q3x_kb_projectors: invalid argument");
            return Q3X_ERR_ARGUMENT;
        }
        const BasisData& B = b->d;
        const KPointData& K = k->d;
        const Lattice& L = B.lat;
        const int npw = static_cast<int>(K.gidx.size());
        KbPlan plan;
        const q3x_status st = kb_plan(B, K, natoms, atom_pos, atom_species, nspecies, nmesh, r_all, rab_all, nbeta,
                                      l_all, cutoff_all, beta_r_all, plan);
        if (st != Q3X_OK) return st;
        if (nproj_out) *nproj_out = plan.nproj;
        if (!beta_out) return Q3X_OK;
        const KbTables& table = *plan.tables;
        const double* kc = plan.kc;
        const int lmax = plan.lmax;
        const int nlm = (lmax + 1) * (lmax + 1);
        std::vector<double> ylm(static_cast<size_t>(npw) * nlm), qabs(npw);
        for (int p = 0; p < npw; ++p) {
            const double* g = B.g[K.gidx[p]].g;
            const double v[3] = {kc[0] + g[0], kc[1] + g[1], kc[2] + g[2]};
            const double v2 = v[0] * v[0] + v[1] * v[1] + v[2] * v[2];
            qabs[p] = std::sqrt(v2) * L.tpa;
            real_ylm(lmax, v, v2, &ylm[static_cast<size_t>(p) * nlm]);
        }
        const double inv_sqrt_vol = 1.0 / std::sqrt(L.volume);
        auto* out = reinterpret_cast<zcplx*>(beta_out);
        std::vector<double> radial(npw);
        std::vector<zcplx> phase(npw);
        int col = 0;
        for (int a = 0; a < natoms; ++a) {
            const int s = plan.species_of_atom[a];
            const double* tau = &plan.tau[3 * static_cast<size_t>(a)];
            const double kdt = plan.kdt[a];
            for (int p = 0; p < npw; ++p) {
                const double* g = B.g[K.gidx[p]].g;
                const double ph = kTwoPi * ((g[0] * tau[0] + g[1] * tau[1] + g[2] * tau[2]) + kdt);
                phase[p] = zcplx(std::cos(ph), -std::sin(ph));
            }
            for (size_t i = 0; i < plan.beta_l[s].size(); ++i) {
                const int l = plan.beta_l[s][i];
                for (int p = 0; p < npw; ++p) radial[p] = interp4(table[s][i], qabs[p]);
                static const zcplx minus_i_pow[4] = {{1, 0}, {0, -1}, {-1, 0}, {0, 1}};
                const zcplx pref = inv_sqrt_vol * minus_i_pow[l % 4];
                for (int mi = 0; mi < 2 * l + 1; ++mi) {
                    const int slot = lm_slot(l, kb_channel_m(mi));
                    zcplx* dst = out + static_cast<size_t>(col) * npw;
                    for (int p = 0; p < npw; ++p)
                        dst[p] = radial[p] * ylm[static_cast<size_t>(p) * nlm + slot] * phase[p] * pref;
                    ++col;
                }
            }
        }
        return Q3X_OK;
    } catch (const std::exception& e) {  // nothing crosses the C ABI as an exception
        set_error(std::string("q3x_kb_projectors: ") + e.what());
        return Q3X_ERR_INTERNAL;
    }
}

// Per-atom D blocks (d_matrix_element, quantum_system/kohn_sham/planewave/norm_conserving/
// init.hpp:84-100, packed as serial.hpp:603-640 packs them).  A species' channels run beta by
// beta with m = 0, +1, -1, ... (the projector order above); two channels couple through
// dion[beta_a][beta_b] exactly when they carry the same (l, m), so within one l the coupling
// block of two betas is dion[a][b] times the (2l+1) identity.  Entries with |D| <= 1e-15 are
// dropped (d_matrix_element_threshold, tolerances.hpp:99).
extern "C" q3x_status q3x_kb_d_blocks(int natoms, const int* atom_species, int nspecies, const int* nbeta,
                                      const int* l_all, const double* dion_all, int* block_sizes, double* d_blocks,
                                      long* packed_len) {
    try {
        if (natoms < 0 || nspecies < 1 || !nbeta || !l_all || !dion_all || (natoms > 0 && !atom_species)) {
            set_error("q3x_kb_d_blocks: invalid argument");
            return Q3X_ERR_ARGUMENT;
        }
        // channel tables per species: (l, m) slot and owning beta
        std::vector<std::vector<int>> slot(nspecies), owner(nspecies);
        std::vector<size_t> dion_off(nspecies);
        size_t lo = 0, doff = 0;
        for (int s = 0; s < nspecies; ++s) {
            if (nbeta[s] < 0) {
                set_error("q3x_kb_d_blocks: negative projector count");
                return Q3X_ERR_ARGUMENT;
            }
            dion_off[s] = doff;
            for (int i = 0; i < nbeta[s]; ++i) {
                const int l = l_all[lo + i];
                if (l < 0 || l > 3) {
                    set_error("q3x_kb_d_blocks: projector l must be 0..3");
                    return Q3X_ERR_ARGUMENT;
                }
                for (int mi = 0; mi < 2 * l + 1; ++mi) {
                    const int m = mi == 0 ? 0 : ((mi & 1) ? (mi + 1) / 2 : -(mi / 2));
                    slot[s].push_back(lm_slot(l, m));
                    owner[s].push_back(i);
                }
            }
            lo += nbeta[s];
            doff += static_cast<size_t>(nbeta[s]) * nbeta[s];
        }
        long total = 0;
        for (int a = 0; a < natoms; ++a) {
            const int s = atom_species[a];
            if (s < 0 || s >= nspecies) {
                set_error("q3x_kb_d_blocks: atom species out of range");
                return Q3X_ERR_ARGUMENT;
            }
            const int p = static_cast<int>(slot[s].size());
            if (block_sizes) block_sizes[a] = p;
            if (d_blocks) {
                double* blk = d_blocks + total;
                std::fill(blk, blk + static_cast<long>(p) * p, 0.0);
                const double* dion = dion_all + dion_off[s];
                for (int col = 0; col < p; ++col)
                    for (int row = 0; row < p; ++row) {
                        if (slot[s][row] != slot[s][col]) continue;
                        const double v = dion[owner[s][row] * nbeta[s] + owner[s][col]];
                        if (std::abs(v) > 1e-15) blk[row + static_cast<long>(col) * p] = v;
                    }
            }
            total += static_cast<long>(p) * p;
        }
        if (packed_len) *packed_len = total;
        return Q3X_OK;
    } catch (const std::exception& e) {  // nothing crosses the C ABI as an exception
        set_error(std::string("q3x_kb_d_blocks: ") + e.what());
        return Q3X_ERR_INTERNAL;
    }
}
