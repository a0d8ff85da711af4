// Internal declarations shared by the host and device translation units of libq3x.
#pragma once

#include <complex>
#include <cstdint>
#include <string>
#include <vector>

#include "../../include/q3x.h"

namespace q3x {

using zcplx = std::complex<double>;

void set_error(const std::string& msg);

// ---------------------------------------------------------------- lattice ----
// Rows a1, a2, a3 in alat units.  Reciprocal rows b_i with b_i . a_j = delta_ij.
struct Lattice {
    double a[3][3];
    double b[3][3];
    double alat;
    double tpa;   // 2 pi / alat
    double tpa2;  // (2 pi / alat)^2
    double volume;
};
Lattice make_lattice(const double rows[9], double alat);

// Smallest 2,3,5-smooth n' >= n divisible by factor (grid_sizer.hpp:173-192).
int smooth_fft_length(int n, int factor);
void fft_box_for_cutoff(const Lattice& L, double ecut_ry, const int factors[3], int dims[3]);

struct GEntry {
    int m[3];
    double g[3];
    double g2;
};

struct BasisData {
    Lattice lat;
    bool gamma = false;  // Gamma trick: half sphere (g_vectors_grid.hpp:56-76)
    double ecutwfc, ecutrho;
    int rho[3], wave[3];
    std::vector<GEntry> g;
};

// Stick decomposition of one k-point's sphere for the pencil FFT.
//   sticks: distinct (m1, m2) pairs, ordered by (column, i2); column = rank of wrapped i1.
//   entries: one per plane wave, stick-major, z ascending within a stick.
struct StickLayout {
    int nst = 0;                    // number of z-sticks
    int ncol = 0;                   // number of distinct i1 columns
    std::vector<int> col_i1;        // ncol: wrapped i1 of each column
    std::vector<int> stick_col;     // nst: column index of stick
    std::vector<int> stick_i2;      // nst: wrapped i2
    std::vector<int> stick_off;     // nst+1: entry range
    std::vector<int> ent_pw;        // npw: plane-wave index (reference order)
    std::vector<int> ent_pos;       // npw: stick*N3 + wrapped i3
};

struct KPointData {
    double k[3];
    std::vector<int> gidx;
    std::vector<double> ekin;
    std::vector<int> fftmap;
    // Gamma trick (k = 0 on a half-sphere basis): box offset of -G per plane wave
    // (serial.hpp:889-904), and the extended list the transforms run over: the npw half-sphere
    // entries, then the mirrors -G_j of j = 1..npw-1 (mirror j at npw + j - 1), with their box
    // offsets and kinetic energies.  sticks index the extended list.
    bool gamma = false;
    std::vector<int> minusmap;
    std::vector<int> ext_fftmap;
    std::vector<double> ext_ekin;
    StickLayout sticks;
};

void build_kpoint(const BasisData& b, const double k[3], KPointData& kp);

}  // namespace q3x

struct q3x_basis_s {
    q3x::BasisData d;
};

struct q3x_kpoint_s {
    const q3x_basis_s* basis;
    q3x::KPointData d;
};
