// What a Kleinman-Bylander projector set of one k-point is built from (kb_projectors.cpp): the
// beta(q) radial tables of the species, the channel layout and the atoms' positions.  The host
// builder (q3x_kb_projectors) and the device builder (q3x_hamiltonian_build_projectors) share it.
#pragma once
#include <memory>
#include <vector>

#include "q3x_internal.h"

namespace q3x {

using KbTables = std::vector<std::vector<std::vector<double>>>;  // [species][beta][q]

struct KbPlan {
    std::shared_ptr<const KbTables> tables;
    int nq = 0;      // q points of every table, spacing 0.01 bohr^-1
    int lmax = 0;    // largest l of any projector
    int nproj = 0;   // columns of the set
    double kc[3] = {0, 0, 0};  // k in cartesian 2 pi / alat units
    std::vector<int> species_of_atom, col0_of_atom;
    std::vector<double> tau;  // natoms x 3, cartesian, alat units
    std::vector<double> kdt;  // natoms: k . tau
    std::vector<std::vector<int>> beta_l;  // per species: l of each beta
};

// Validates the species arrays and fills the plan (tables from the shared cache).  The
// arguments are those of q3x_kb_projectors.
q3x_status kb_plan(const BasisData& B, const KPointData& K, int natoms, const double* atom_pos,
                   const int* atom_species, int nspecies, const int* nmesh, const double* r_all,
                   const double* rab_all, const int* nbeta, const int* l_all, const int* cutoff_all,
                   const double* beta_r_all, KbPlan& plan);

// Channel order within one beta of angular momentum l: m = 0, +1, -1, +2, -2, ...
// (harmonics.hpp:40-69); slot of (l, m) in a real Y_lm array.
inline int kb_channel_m(int mi) { return mi == 0 ? 0 : ((mi & 1) ? (mi + 1) / 2 : -(mi / 2)); }
inline int kb_lm_slot(int l, int m) { return m == 0 ? l * l : (m > 0 ? l * l + 2 * m - 1 : l * l - 2 * m); }

}  // namespace q3x
