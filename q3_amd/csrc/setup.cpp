// Host-side setup of the path: lattice, FFT box sizing, G sphere, per-k plane-wave lists,
// FFT-box maps and the z-stick layout the pencil FFT kernels run on.
//
// The G sphere and per-k lists must agree bit for bit with the reference, so the
// floating-point expressions below keep the reference's operation order and the sorts use
// std::sort with the reference's comparators on the reference's input order.  FP
// contraction is off for this file (no FMA), as in an x86-64 build of the reference.
#pragma clang fp contract(off)

#include <algorithm>
#include <cmath>
#include <map>
#include <numeric>
#include <stdexcept>

#include "q3x_internal.h"

namespace q3x {

namespace {
constexpr double kPi = 3.14159265358979323846;
constexpr double kTwoPi = 2.0 * kPi;
constexpr double kCutoffSlack = 1e-8;  // energy_cutoff_rounding_tolerance (tolerances.hpp:129)
}  // namespace

// Crystal::compute_derived_quantities (crystal.hpp:156-177) with Matrix3x3::determinant and
// inverse (matrix_3x3.hpp:51-71): B = (A^-1)^T.
Lattice make_lattice(const double rows[9], double alat) {
    Lattice L{};
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) L.a[i][j] = rows[3 * i + j];
    const double(&r)[3][3] = L.a;
    const double det = r[0][0] * (r[1][1] * r[2][2] - r[1][2] * r[2][1]) -
                       r[0][1] * (r[1][0] * r[2][2] - r[1][2] * r[2][0]) +
                       r[0][2] * (r[1][0] * r[2][1] - r[1][1] * r[2][0]);
    double inv[3][3];
    inv[0][0] = (r[1][1] * r[2][2] - r[1][2] * r[2][1]) / det;
    inv[0][1] = (r[0][2] * r[2][1] - r[0][1] * r[2][2]) / det;
    inv[0][2] = (r[0][1] * r[1][2] - r[0][2] * r[1][1]) / det;
    inv[1][0] = (r[1][2] * r[2][0] - r[1][0] * r[2][2]) / det;
    inv[1][1] = (r[0][0] * r[2][2] - r[0][2] * r[2][0]) / det;
    inv[1][2] = (r[0][2] * r[1][0] - r[0][0] * r[1][2]) / det;
    inv[2][0] = (r[1][0] * r[2][1] - r[1][1] * r[2][0]) / det;
    inv[2][1] = (r[0][1] * r[2][0] - r[0][0] * r[2][1]) / det;
    inv[2][2] = (r[0][0] * r[1][1] - r[0][1] * r[1][0]) / det;
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) L.b[i][j] = inv[j][i];
    L.alat = alat;
    L.tpa = kTwoPi / alat;
    L.tpa2 = L.tpa * L.tpa;
    L.volume = std::abs(det) * alat * alat * alat;
    return L;
}

int smooth_fft_length(int n, int factor) {
    for (int c = n;; ++c) {
        if (factor > 1 && c % factor) continue;
        int rest = c;
        for (int p : {2, 3, 5})
            while (rest % p == 0) rest /= p;
        if (rest == 1) return c;
    }
}

static inline void g_of_miller(const Lattice& L, int i1, int i2, int i3, double g[3]) {
    for (int d = 0; d < 3; ++d) g[d] = i1 * L.b[0][d] + i2 * L.b[1][d] + i3 * L.b[2][d];
}
static inline double norm2(const double v[3]) { return v[0] * v[0] + v[1] * v[1] + v[2] * v[2]; }

// FFTGridSizer::calculate_grid_dimensions (grid_sizer.hpp:101-160).
void fft_box_for_cutoff(const Lattice& L, double ecut_ry, const int factors[3], int dims[3]) {
    const double gmax2 = ecut_ry / L.tpa2;
    const double gmax = std::sqrt(gmax2);
    int reach[3];
    for (int d = 0; d < 3; ++d) reach[d] = static_cast<int>(std::floor(gmax * std::sqrt(norm2(L.a[d])))) + 1;
    int extent[3] = {0, 0, 0};
    for (int i3 = -reach[2]; i3 <= reach[2]; ++i3)
        for (int i2 = -reach[1]; i2 <= reach[1]; ++i2)
            for (int i1 = -reach[0]; i1 <= reach[0]; ++i1) {
                double g[3];
                g_of_miller(L, i1, i2, i3, g);
                if (norm2(g) < gmax2 + kCutoffSlack) {
                    extent[0] = std::max(extent[0], std::abs(i1));
                    extent[1] = std::max(extent[1], std::abs(i2));
                    extent[2] = std::max(extent[2], std::abs(i3));
                }
            }
    for (int d = 0; d < 3; ++d) dims[d] = smooth_fft_length(2 * extent[d] + 1, factors ? factors[d] : 1);
}

// G sphere of the density grid (g_vectors_grid.hpp:23-104, serial full sphere): i1, i2, i3
// nested in that order, |G|^2 <= ecutrho/(2pi/a)^2 + 1e-8, then std::sort by |G|^2.
static void build_sphere(BasisData& B) {
    const Lattice& L = B.lat;
    const double cut = B.ecutrho / L.tpa2 + kCutoffSlack;
    const int h1 = B.rho[0] / 2, h2 = B.rho[1] / 2, h3 = B.rho[2] / 2;
    B.g.clear();
    B.g.reserve(static_cast<size_t>(B.rho[0]) * B.rho[1] * B.rho[2] / 4);
    for (int i1 = -h1; i1 <= h1; ++i1)
        for (int i2 = -h2; i2 <= h2; ++i2)
            for (int i3 = -h3; i3 <= h3; ++i3) {
                if (B.gamma) {  // g_vectors_grid.hpp:69-76: keep i3 > 0, (0, i2 > 0), (0, 0, i1 >= 0)
                    if (i3 < 0) continue;
                    if (i3 == 0 && i2 < 0) continue;
                    if (i3 == 0 && i2 == 0 && i1 < 0) continue;
                }
                GEntry e;
                g_of_miller(L, i1, i2, i3, e.g);
                e.g2 = norm2(e.g);
                if (e.g2 <= cut) {
                    e.m[0] = i1;
                    e.m[1] = i2;
                    e.m[2] = i3;
                    B.g.push_back(e);
                }
            }
    std::sort(B.g.begin(), B.g.end(), [](const GEntry& x, const GEntry& y) { return x.g2 < y.g2; });
}

static int wrap(int m, int n) { return m >= 0 ? m : m + n; }

// Stick layout for the pencil FFT.  Permutation-only bookkeeping: values are never reordered
// in HBM, the kernels gather/scatter through ent_pw.
// millers: 3 ints per plane wave of the list the transforms run over; fftmap: its box offsets
// (-1 outside the box).
static void build_sticks(const int wave[3], const std::vector<int>& millers, const std::vector<int>& fftmap,
                         StickLayout& S) {
    const int n1 = wave[0], n2 = wave[1], n3 = wave[2];
    const int npw = static_cast<int>(fftmap.size());
    // distinct columns (wrapped i1) present
    std::vector<char> has_col(n1, 0);
    for (int p = 0; p < npw; ++p) {
        if (fftmap[p] < 0) continue;
        has_col[wrap(millers[3 * p], n1)] = 1;
    }
    std::vector<int> col_of_i1(n1, -1);
    S.col_i1.clear();
    for (int i = 0; i < n1; ++i)
        if (has_col[i]) {
            col_of_i1[i] = static_cast<int>(S.col_i1.size());
            S.col_i1.push_back(i);
        }
    S.ncol = static_cast<int>(S.col_i1.size());
    // sticks keyed by (column, wrapped i2)
    std::vector<int> stick_id(static_cast<size_t>(S.ncol) * n2, -1);
    std::vector<int> pw_stick(npw, -1), pw_z(npw, -1);
    for (int p = 0; p < npw; ++p) {
        if (fftmap[p] < 0) continue;
        const int* m = &millers[3 * p];
        int c = col_of_i1[wrap(m[0], n1)];
        int key = c * n2 + wrap(m[1], n2);
        stick_id[key] = 0;
        pw_stick[p] = key;
        pw_z[p] = wrap(m[2], n3);
    }
    S.stick_col.clear();
    S.stick_i2.clear();
    for (int key = 0; key < S.ncol * n2; ++key)
        if (stick_id[key] == 0) {
            stick_id[key] = static_cast<int>(S.stick_col.size());
            S.stick_col.push_back(key / n2);
            S.stick_i2.push_back(key % n2);
        }
    S.nst = static_cast<int>(S.stick_col.size());
    // entries: counting sort by stick, then by z within a stick
    std::vector<int> count(S.nst + 1, 0);
    for (int p = 0; p < npw; ++p)
        if (pw_stick[p] >= 0) count[stick_id[pw_stick[p]] + 1]++;
    for (int s = 0; s < S.nst; ++s) count[s + 1] += count[s];
    S.stick_off = count;
    int nent = count[S.nst];
    S.ent_pw.assign(nent, -1);
    S.ent_pos.assign(nent, -1);
    std::vector<int> fill(count.begin(), count.end() - 1);
    for (int p = 0; p < npw; ++p) {
        if (pw_stick[p] < 0) continue;
        int s = stick_id[pw_stick[p]];
        S.ent_pw[fill[s]++] = p;
    }
    for (int s = 0; s < S.nst; ++s) {
        auto first = S.ent_pw.begin() + S.stick_off[s];
        auto last = S.ent_pw.begin() + S.stick_off[s + 1];
        std::sort(first, last, [&](int x, int y) { return pw_z[x] < pw_z[y]; });
        for (int j = S.stick_off[s]; j < S.stick_off[s + 1]; ++j) S.ent_pos[j] = s * n3 + pw_z[S.ent_pw[j]];
    }
}

// PlaneWaveSelector::select (plane_wave_selector.hpp:41-109) and the G -> FFT-box map
// (planewave_potential_utils.hpp:141-164, fft_grid_serial.hpp:301-306).
void build_kpoint(const BasisData& B, const double k[3], KPointData& kp) {
    const Lattice& L = B.lat;
    for (int d = 0; d < 3; ++d) kp.k[d] = k[d];
    double kc[3];
    for (int d = 0; d < 3; ++d) kc[d] = k[0] * L.b[0][d] + k[1] * L.b[1][d] + k[2] * L.b[2][d];
    const double cut = B.ecutwfc / L.tpa2 + kCutoffSlack;
    std::vector<int> sel;
    std::vector<double> kin;
    for (int i = 0; i < static_cast<int>(B.g.size()); ++i) {
        const double v[3] = {kc[0] + B.g[i].g[0], kc[1] + B.g[i].g[1], kc[2] + B.g[i].g[2]};
        const double q2 = norm2(v);
        if (q2 <= cut) {
            sel.push_back(i);
            kin.push_back(q2 * L.tpa2);
        }
    }
    std::vector<int> order(sel.size());
    std::iota(order.begin(), order.end(), 0);
    std::sort(order.begin(), order.end(), [&kin](int x, int y) { return kin[x] < kin[y]; });
    const int npw = static_cast<int>(sel.size());
    kp.gidx.resize(npw);
    kp.ekin.resize(npw);
    kp.fftmap.resize(npw);
    const int n1 = B.wave[0], n2 = B.wave[1], n3 = B.wave[2];
    auto box = [n1, n2, n3](int m1, int m2, int m3) {
        if (std::abs(m1) <= n1 / 2 && std::abs(m2) <= n2 / 2 && std::abs(m3) <= n3 / 2)
            return (wrap(m1, n1) * n2 + wrap(m2, n2)) * n3 + wrap(m3, n3);
        return -1;
    };
    std::vector<int> millers(3 * static_cast<size_t>(npw));
    for (int p = 0; p < npw; ++p) {
        kp.gidx[p] = sel[order[p]];
        kp.ekin[p] = kin[order[p]];
        const int* m = B.g[kp.gidx[p]].m;
        kp.fftmap[p] = box(m[0], m[1], m[2]);
        for (int d = 0; d < 3; ++d) millers[3 * p + d] = m[d];
    }
    kp.gamma = B.gamma;
    kp.minusmap.clear();
    kp.ext_fftmap.clear();
    kp.ext_ekin.clear();
    if (!B.gamma) {
        build_sticks(B.wave, millers, kp.fftmap, kp.sticks);
        return;
    }
    // Gamma trick: -G map (serial.hpp:889-904) and the extended list [half sphere | mirrors of 1..]
    kp.minusmap.resize(npw);
    for (int p = 0; p < npw; ++p) kp.minusmap[p] = box(-millers[3 * p], -millers[3 * p + 1], -millers[3 * p + 2]);
    const int next = npw > 0 ? 2 * npw - 1 : 0;
    std::vector<int> ext_m(3 * static_cast<size_t>(next));
    kp.ext_fftmap.resize(next);
    kp.ext_ekin.resize(next);
    for (int p = 0; p < npw; ++p) {
        for (int d = 0; d < 3; ++d) ext_m[3 * p + d] = millers[3 * p + d];
        kp.ext_fftmap[p] = kp.fftmap[p];
        kp.ext_ekin[p] = kp.ekin[p];
        if (p == 0) continue;
        const int q = npw + p - 1;
        for (int d = 0; d < 3; ++d) ext_m[3 * q + d] = -millers[3 * p + d];
        kp.ext_fftmap[q] = kp.minusmap[p];
        kp.ext_ekin[q] = kp.ekin[p];
    }
    build_sticks(B.wave, ext_m, kp.ext_fftmap, kp.sticks);
}

}  // namespace q3x

using namespace q3x;

// ------------------------------------------------------------------ C ABI ----
extern "C" {

q3x_status q3x_basis_create(const double lattice_rows[9], double alat_bohr, double ecutwfc_ry,
                            double ecutrho_ry, int use_exact_wavefunction_grid,
                            const int fft_factors[3], q3x_basis* out) {
    return q3x_basis_create_gamma(lattice_rows, alat_bohr, ecutwfc_ry, ecutrho_ry, use_exact_wavefunction_grid,
                                  fft_factors, 0, out);
}

q3x_status q3x_basis_create_gamma(const double lattice_rows[9], double alat_bohr, double ecutwfc_ry,
                                  double ecutrho_ry, int use_exact_wavefunction_grid, const int fft_factors[3],
                                  int gamma_only, q3x_basis* out) {
    if (!lattice_rows || !out || !(alat_bohr > 0) || !(ecutwfc_ry > 0)) {
        set_error("q3x_basis_create: invalid argument");
        return Q3X_ERR_ARGUMENT;
    }
    try {
        auto* b = new q3x_basis_s;
        BasisData& B = b->d;
        B.gamma = gamma_only != 0;
        B.lat = make_lattice(lattice_rows, alat_bohr);
        if (!(B.lat.volume > 1e-10)) {
            delete b;
            set_error("q3x_basis_create: singular lattice");
            return Q3X_ERR_ARGUMENT;
        }
        B.ecutwfc = ecutwfc_ry;
        B.ecutrho = ecutrho_ry > 0.0 ? ecutrho_ry : 4.0 * ecutwfc_ry;
        int one[3] = {1, 1, 1};
        const int* f = fft_factors ? fft_factors : one;
        for (int d = 0; d < 3; ++d)
            if (f[d] < 1 || f[d] > 6) {
                delete b;
                set_error("q3x_basis_create: fft factor outside 1..6");
                return Q3X_ERR_ARGUMENT;
            }
        fft_box_for_cutoff(B.lat, B.ecutrho, f, B.rho);
        // pair.hpp:86-97: wave grid for min(4 or 2 x ecutwfc, ecutrho)
        double smooth = std::min((use_exact_wavefunction_grid ? 4.0 : 2.0) * B.ecutwfc, B.ecutrho);
        fft_box_for_cutoff(B.lat, smooth, f, B.wave);
        build_sphere(B);
        *out = b;
        return Q3X_OK;
    } catch (const std::exception& e) {
        set_error(std::string("q3x_basis_create: ") + e.what());
        return Q3X_ERR_INTERNAL;
    }
}

q3x_status q3x_basis_info(q3x_basis b, int rho_dims[3], int wave_dims[3], int* ng, double* volume,
                          double* tpa2) {
    if (!b) {
        set_error("q3x_basis_info: null basis");
        return Q3X_ERR_ARGUMENT;
    }
    for (int d = 0; d < 3; ++d) {
        if (rho_dims) rho_dims[d] = b->d.rho[d];
        if (wave_dims) wave_dims[d] = b->d.wave[d];
    }
    if (ng) *ng = static_cast<int>(b->d.g.size());
    if (volume) *volume = b->d.lat.volume;
    if (tpa2) *tpa2 = b->d.lat.tpa2;
    return Q3X_OK;
}

q3x_status q3x_basis_is_gamma_only(q3x_basis b, int* gamma_only) {
    if (!b || !gamma_only) {
        set_error("q3x_basis_is_gamma_only: invalid argument");
        return Q3X_ERR_ARGUMENT;
    }
    *gamma_only = b->d.gamma ? 1 : 0;
    return Q3X_OK;
}

q3x_status q3x_basis_gvectors(q3x_basis b, int* miller, double* g_cart, double* g2) {
    if (!b) {
        set_error("q3x_basis_gvectors: null basis");
        return Q3X_ERR_ARGUMENT;
    }
    const auto& G = b->d.g;
    for (size_t i = 0; i < G.size(); ++i) {
        for (int d = 0; d < 3; ++d) {
            if (miller) miller[3 * i + d] = G[i].m[d];
            if (g_cart) g_cart[3 * i + d] = G[i].g[d];
        }
        if (g2) g2[i] = G[i].g2;
    }
    return Q3X_OK;
}

q3x_status q3x_basis_destroy(q3x_basis b) {
    delete b;
    return Q3X_OK;
}

q3x_status q3x_kpoint_create(q3x_basis b, const double k_crystal[3], q3x_kpoint* out, int* npw) {
    if (!b || !k_crystal || !out) {
        set_error("q3x_kpoint_create: invalid argument");
        return Q3X_ERR_ARGUMENT;
    }
    if (b->d.gamma && (k_crystal[0] != 0.0 || k_crystal[1] != 0.0 || k_crystal[2] != 0.0)) {
        set_error("q3x_kpoint_create: a Gamma-trick (half-sphere) basis has only the k-point (0, 0, 0)");
        return Q3X_ERR_ARGUMENT;
    }
    try {
        auto* k = new q3x_kpoint_s;
        k->basis = b;
        build_kpoint(b->d, k_crystal, k->d);
        if (npw) *npw = static_cast<int>(k->d.gidx.size());
        *out = k;
        return Q3X_OK;
    } catch (const std::exception& e) {
        set_error(std::string("q3x_kpoint_create: ") + e.what());
        return Q3X_ERR_INTERNAL;
    }
}

q3x_status q3x_kpoint_maps(q3x_kpoint k, int* g_index, double* kinetic_ry, int* fft_map) {
    if (!k) {
        set_error("q3x_kpoint_maps: null k-point");
        return Q3X_ERR_ARGUMENT;
    }
    const auto& d = k->d;
    for (size_t i = 0; i < d.gidx.size(); ++i) {
        if (g_index) g_index[i] = d.gidx[i];
        if (kinetic_ry) kinetic_ry[i] = d.ekin[i];
        if (fft_map) fft_map[i] = d.fftmap[i];
    }
    return Q3X_OK;
}

q3x_status q3x_kpoint_minus_map(q3x_kpoint k, int* minus_fft_map) {
    if (!k || !minus_fft_map) {
        set_error("q3x_kpoint_minus_map: invalid argument");
        return Q3X_ERR_ARGUMENT;
    }
    if (!k->d.gamma) {
        set_error("q3x_kpoint_minus_map: the k-point is not on a Gamma-trick basis");
        return Q3X_ERR_STATE;
    }
    std::copy(k->d.minusmap.begin(), k->d.minusmap.end(), minus_fft_map);
    return Q3X_OK;
}

q3x_status q3x_kpoint_random_psi(q3x_kpoint k, int nb, unsigned long long seed, double* psi_out) {
    if (!k || nb < 0 || (!psi_out && nb > 0)) {
        set_error("q3x_kpoint_random_psi: invalid argument");
        return Q3X_ERR_ARGUMENT;
    }
    const std::vector<double>& ek = k->d.ekin;
    const size_t npw = ek.size();
    // entry (g, b), part r: splitmix64 of (seed, 2 (b npw + g) + r) -> uniform in [-1, 1)
    auto mix = [](unsigned long long x) {
        x += 0x9E3779B97F4A7C15ull;
        x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
        x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
        return x ^ (x >> 31);
    };
    const unsigned long long s0 = mix(seed);
    const double scale = 1.0 / 4503599627370496.0;  // 2^-52
    for (int b = 0; b < nb; ++b)
        for (size_t g = 0; g < npw; ++g) {
            const double damp = 1.0 / (1.0 + ek[g]);
            const size_t o = 2 * (static_cast<size_t>(b) * npw + g);
            const double re = static_cast<double>(mix(s0 + o) >> 11) * scale - 1.0;
            const double im = static_cast<double>(mix(s0 + o + 1) >> 11) * scale - 1.0;
            psi_out[o] = re * damp;
            // a Gamma-trick band is a real function: its G = 0 coefficient is real
            psi_out[o + 1] = k->d.gamma && g == 0 ? 0.0 : im * damp;
        }
    return Q3X_OK;
}

q3x_status q3x_kpoint_destroy(q3x_kpoint k) {
    delete k;
    return Q3X_OK;
}

}  // extern "C"
