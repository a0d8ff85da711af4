/*
 * q3x.h -- C ABI of the MI355X-native H|psi> + block-Davidson library (libq3x.so).
 *
 * The library replaces the plane-wave Hamiltonian-application and Davidson path of
 * camsai/q3.  Each entry point names the reference interface it stands in for
 * (paths relative to the camsai/q3 source tree).
 *
 * Conventions (the reference's, kept so the library drops in):
 *   - complex numbers are interleaved (re, im) doubles, i.e. std::complex<double>;
 *   - band blocks are column-major, one band per column, leading dimension npw
 *     (src/layer_3_representation/quantum_system/kohn_sham/planewave/init.hpp:120-138);
 *   - projector sets are npw x nproj, one projector after another
 *     (src/layer_3_representation/pseudopotential/planewave/kleinman_bylander_projector_set.hpp:20-23);
 *   - energies are in Rydberg; lengths in bohr; lattice vectors in units of alat;
 *   - every call returns a q3x_status; on failure q3x_last_error() holds a message
 *     (thread-local).  No C++ exception crosses the ABI.
 *   - handles are not thread-safe; one context per host thread (the reference's
 *     solver is not thread-safe either: src/bindings/q3_core_api.hpp:12-14).
 */
#ifndef Q3X_H
#define Q3X_H

#ifdef __cplusplus
extern "C" {
#endif

typedef int q3x_status;
enum {
    Q3X_OK = 0,
    Q3X_ERR_ARGUMENT = 1,
    Q3X_ERR_DEVICE = 2,
    Q3X_ERR_STATE = 3,    /* e.g. apply before prepare_kpoint (serial.hpp:487-492 logic_error) */
    Q3X_ERR_NUMERIC = 4,  /* eigensolver / Cholesky breakdown that recovery could not fix */
    Q3X_ERR_INTERNAL = 5
};

/* precision of the H-apply / Davidson products */
enum {
    Q3X_FP64 = 0,  /* everything complex128 */
    Q3X_MIXED = 1  /* FFT, V multiply and projector GEMMs in complex64 with fp32 accumulation,
                      and in Davidson at tolerances of 1e-6 Ry and above also the products that
                      form basis vectors; Gram products, Cholesky, Rayleigh-Ritz and the final
                      residual sweep in fp64 */
};

typedef struct q3x_context_s* q3x_context;
typedef struct q3x_basis_s* q3x_basis;
typedef struct q3x_kpoint_s* q3x_kpoint;
typedef struct q3x_hamiltonian_s* q3x_hamiltonian;

const char* q3x_last_error(void);
const char* q3x_version(void);

/* ---- device context (stands in for GPUDeviceBackend,
 *      src/layer_0_core/execution/gpu/device/init.hpp; explicit hipMalloc, no managed memory) */
q3x_status q3x_context_create(int device, q3x_context* out);
q3x_status q3x_context_destroy(q3x_context ctx);
q3x_status q3x_context_synchronize(q3x_context ctx);

/* ---- plane-wave basis (PlanewaveBasis::generate,
 *      src/layer_3_representation/basis/planewave/init.hpp:60-109, with
 *      PlanewaveFFTPair::initialize .../internal/fft/pair.hpp:24-100 and
 *      GVectorsGrid::generate .../internal/g_vectors_grid.hpp:23-165).
 *      fft_factors replaces spglib's compute_fft_factors (grid_sizer.hpp:83-94): the caller
 *      passes the divisibility factors.  ecutrho_ry <= 0 means 4 x ecutwfc (init.hpp:70-72).
 *      Host-only: needs no device. */
q3x_status q3x_basis_create(const double lattice_rows[9], double alat_bohr, double ecutwfc_ry,
                            double ecutrho_ry, int use_exact_wavefunction_grid,
                            const int fft_factors[3], q3x_basis* out);
/* The same with the Gamma trick (gamma_only != 0): the G list keeps one of each +-G pair, the half
 * sphere i3 > 0, or i3 = 0 and i2 > 0, or i3 = i2 = 0 and i1 >= 0 (GVectorsGrid::generate with
 * is_gamma_only, g_vectors_grid.hpp:56-76).  The reference switches it on for a 1x1x1 unshifted
 * k grid (iterator/kohn_sham/planewave/init.hpp:585-610).  Its only k-point is (0, 0, 0); band
 * blocks on it hold the half-sphere coefficients of real functions (psi(-G) = conj(psi(G)),
 * G = 0 real and first), and every product over plane waves is the full-sphere real one. */
q3x_status q3x_basis_create_gamma(const double lattice_rows[9], double alat_bohr, double ecutwfc_ry,
                                  double ecutrho_ry, int use_exact_wavefunction_grid,
                                  const int fft_factors[3], int gamma_only, q3x_basis* out);
q3x_status q3x_basis_is_gamma_only(q3x_basis b, int* gamma_only);
q3x_status q3x_basis_info(q3x_basis b, int rho_dims[3], int wave_dims[3], int* number_of_g_vectors,
                          double* cell_volume, double* two_pi_over_alat_squared);
/* miller: ng x 3 ints; g_cart: ng x 3 (2pi/alat units); g2: ng (2pi/alat)^2 units; any may be NULL */
q3x_status q3x_basis_gvectors(q3x_basis b, int* miller, double* g_cart, double* g2);
q3x_status q3x_basis_destroy(q3x_basis b);

/* ---- k-point plane-wave list (PlaneWaveSelector::select,
 *      .../internal/kpoint/plane_wave_selector.hpp:41-109; FFT map of
 *      compute_g_to_fft_index_mapping, hamiltonian_applicator/planewave_potential_utils.hpp:141-164).
 *      Host-only. */
q3x_status q3x_kpoint_create(q3x_basis b, const double k_crystal[3], q3x_kpoint* out, int* npw);
/* g_index: npw ints into the basis G list; kinetic: |k+G|^2 in Ry; fft_map: wave-grid offsets */
q3x_status q3x_kpoint_maps(q3x_kpoint k, int* g_index, double* kinetic_ry, int* fft_map);
/* Gamma trick: wave-grid offsets of -G for each plane wave (the -G map of
 * hamiltonian_applicator/serial.hpp:889-904); Q3X_ERR_STATE on a full-sphere k-point. */
q3x_status q3x_kpoint_minus_map(q3x_kpoint k, int* minus_fft_map);
/* Starting block for a solve: npw x nb complex128, column-major, real and imaginary parts
 * uniform in [-1, 1) (a splitmix64 hash of seed and the entry's position) damped by
 * 1 / (1 + |k+G|^2 [Ry]).  The band-structure server seeds k-point i of a request with seed i,
 * so a request's solve is the same alone and among others.  Host-only; no counterpart in the
 * reference's interface (its random starting wavefunctions come from its own generator). */
q3x_status q3x_kpoint_random_psi(q3x_kpoint k, int nb, unsigned long long seed, double* psi_out);
q3x_status q3x_kpoint_destroy(q3x_kpoint k);

/* ---- Kleinman-Bylander projector set at one k-point (compute_kleinman_bylander_projectors,
 *      src/layer_3_representation/pseudopotential/planewave/internal/projector.hpp:42-199).
 *      Species arrays are concatenated over species: nmesh[s] radial points (r, rab),
 *      nbeta[s] projectors with angular momentum l and cutoff index, radial values r*beta(r)
 *      (UPF convention), nbeta[s] x nmesh[s] per species.  atom_pos in crystal coordinates.
 *      beta_out: npw x nproj complex (caller-sized; *nproj returns the count).  Host-only. */
q3x_status q3x_kb_projectors(q3x_basis b, q3x_kpoint k, int natoms, const double* atom_pos,
                             const int* atom_species, int nspecies, const int* nmesh,
                             const double* r_all, const double* rab_all, const int* nbeta,
                             const int* l_all, const int* cutoff_all, const double* beta_r_all,
                             double* beta_out, int* nproj);

/* ---- per-atom D blocks of the nonlocal term, in the layout q3x_hamiltonian_set_projectors
 *      takes (NormConservingPlanewaveKohnShamQuantumSystem::d_matrix_element,
 *      src/layer_3_representation/quantum_system/kohn_sham/planewave/norm_conserving/init.hpp:84-100,
 *      packed per atom as hamiltonian_applicator/serial.hpp:603-640 packs them).  Channels of a
 *      species follow the projector order of q3x_kb_projectors (beta, then m = 0, +1, -1, ...);
 *      D_ab = dion[beta_a][beta_b] when channels a and b have the same (l, m), else 0, dropped when
 *      |D| <= 1e-15 (tolerances.hpp:99).  dion_all: per species nbeta[s] x nbeta[s] in Ry,
 *      row-major, concatenated.  block_sizes (natoms) and d_blocks (column-major blocks packed in
 *      atom order) may be NULL; *packed_len returns the packed length.  Host-only. */
q3x_status q3x_kb_d_blocks(int natoms, const int* atom_species, int nspecies, const int* nbeta,
                           const int* l_all, const double* dion_all, int* block_sizes, double* d_blocks,
                           long* packed_len);

/* ---- Hamiltonian applicator (PlanewaveHamiltonianApplicator,
 *      src/layer_4_computation/diagonalizer/planewave/internal/hamiltonian_applicator/init.hpp,
 *      serial semantics of .../serial.hpp:64-155) */
q3x_status q3x_hamiltonian_create(q3x_context ctx, q3x_basis b, q3x_hamiltonian* out);
q3x_status q3x_hamiltonian_destroy(q3x_hamiltonian h);
/* update_potential (serial.hpp:64-78): V(r) in Ry, C order (i1,i2,i3), on the density grid
 * (n = its n1*n2*n3 points, the reference's real_space_potential).  When the wavefunction grid is
 * smaller, V is restricted to it in G space on the device (compute_smooth_grid_potential,
 * serial.hpp:852-872, and planewave_potential_utils.hpp:33-130: forward FFT / N_dense, keep the
 * smooth box's G, backward FFT, real part); when the grids coincide V is taken as is.  V already
 * on the wavefunction grid (n = its point count) is accepted too.  The library keeps V_smooth/N,
 * as the reference does. */
q3x_status q3x_hamiltonian_set_potential(q3x_hamiltonian h, const double* v_r, long n);
/* prepare_kpoint (serial.hpp:84-103) */
q3x_status q3x_hamiltonian_prepare_kpoint(q3x_hamiltonian h, q3x_kpoint k);
/* projector set of the prepared k-point + per-atom D blocks (serial.hpp:587-628):
 * beta npw x nproj complex; block_sizes per atom; d_blocks real, column-major blocks packed in
 * atom order (Ry).  |D| <= 1e-15 is dropped as the reference does (tolerances.hpp:99). */
q3x_status q3x_hamiltonian_set_projectors(q3x_hamiltonian h, const double* beta, int nproj,
                                          int natoms, const int* block_sizes,
                                          const double* d_blocks);
/* The same projector set built on the device (GPUProjectorSetBuilder::build,
 * hamiltonian_applicator/projector_set_builder_gpu.hpp:89-300): the species arrays are those of
 * q3x_kb_projectors, the D blocks those of q3x_hamiltonian_set_projectors.  Only the beta(q)
 * radial tables (host, cached per species set and cutoff), the k-point's G list and the atoms go
 * to the device; form factors, structure factors and their products are formed there, in
 * complex128 and complex64.  Equal to q3x_kb_projectors + q3x_hamiltonian_set_projectors to
 * rounding (1e-13 relative). */
q3x_status q3x_hamiltonian_build_projectors(q3x_hamiltonian h, int natoms, const double* atom_pos,
                                            const int* atom_species, int nspecies, const int* nmesh,
                                            const double* r_all, const double* rab_all, const int* nbeta,
                                            const int* l_all, const int* cutoff_all,
                                            const double* beta_r_all, const int* block_sizes,
                                            const double* d_blocks);
/* q3x_kpoint_random_psi of the prepared k-point written on the device (d_psi: npw x nb
 * complex128), bit for bit the host generator's block; stream-ordered.  No counterpart in the
 * reference's interface. */
q3x_status q3x_hamiltonian_random_psi(q3x_hamiltonian h, int nb, unsigned long long seed, void* d_psi);
/* apply_batch (serial.hpp:110-155): host buffers psi / hpsi, npw x nvec complex. */
q3x_status q3x_hamiltonian_apply(q3x_hamiltonian h, const double* psi, double* hpsi, int nvec,
                                 int precision);
/* same on device buffers (complex128, ld npw) already resident in HBM; stream-ordered on the
 * context stream, returns after queueing (call q3x_context_synchronize to wait). */
q3x_status q3x_hamiltonian_apply_device(q3x_hamiltonian h, const void* d_psi, void* d_hpsi,
                                        int nvec, int precision);
/* Charge density of the prepared k-point's bands (PlanewaveChargeDensityAccumulator::
 * accumulate_kpoint_state, src/layer_4_computation/charge_density/planewave/internal/
 * accumulator.hpp:159-280, with the per-band kernel of accumulator_gpu.hpp:267):
 * rho(r) += sum_b weights[b] |psi_b(r)|^2 on the wavefunction box, psi_b(r) = sum_G c_b(G)
 * e^{i(k+G).r} (the unnormalised backward transform).  d_psi: npw x nb complex128 in HBM (with the
 * Gamma trick, the half-sphere blocks: pairs of real bands share one transform); weights: nb
 * host doubles (w_k f_nk / Omega for the reference's rho); d_rho: n1 n2 n3 doubles in HBM,
 * C order (i1, i2, i3), added to (the caller clears it).  Stream-ordered.  FP64.  Needs a box the
 * two-level kernels cover with every plane wave inside it (Q3X_ERR_ARGUMENT otherwise). */
q3x_status q3x_hamiltonian_accumulate_density(q3x_hamiltonian h, const void* d_psi, int nb, const double* weights,
                                              void* d_rho);
/* No counterpart in the reference.  For callers that cannot allocate device memory (code under
 * stream capture, serving loops that forbid hipMalloc): the product of
 * q3x_hamiltonian_apply_device with every scratch buffer (stick buffer, plane slots, projector
 * products, split-K workspace) in one fixed device array of q3x_noalloc_scratch_bytes() bytes
 * (64 MiB) per GPU, held until the process ends.  Bands run in chunks that fit it; one band
 * that does not fit is refused with Q3X_ERR_ARGUMENT.  Calls no hipMalloc/hipFree.  Every call
 * on one GPU shares that GPU's array: calls that overlap on one GPU (two contexts or threads)
 * race.  Nothing in the array needs to start at zero. */
q3x_status q3x_hamiltonian_apply_noalloc(q3x_hamiltonian h, const void* d_psi, void* d_hpsi,
                                         int nvec, int precision);
unsigned long q3x_noalloc_scratch_bytes(void);
/* complex GEMM on device buffers, stream-ordered (stands in for GPUBLAS::gemm,
 * src/layer_0_core/math/linear_algebra/gpu/blas/init.hpp, on MFMA):
 *   C (m x n) = alpha op(A) B + beta C,  transa 'N' (A m x k) or 'C' (A k x m, op = A^H),
 *   alpha, beta real; Q3X_FP64: complex128 throughout; Q3X_MIXED: complex128 operands rounded to
 *   complex64 for the products, fp32 accumulation, complex128 C. */
q3x_status q3x_zgemm(q3x_context ctx, char transa, int m, int n, int k, double alpha, const void* A,
                     long lda, const void* B, long ldb, double beta, void* C, long ldc, int precision);
/* device scratch helpers for callers that keep psi resident */
q3x_status q3x_device_alloc(q3x_context ctx, unsigned long bytes, void** out);
q3x_status q3x_device_free(q3x_context ctx, void* p);
/* stream-ordered device-to-device copy by a plain streaming kernel (bytes multiple of 16);
 * also the reference kernel for calibrating HBM byte counters */
q3x_status q3x_device_copy(q3x_context ctx, void* dst, const void* src, unsigned long bytes);
q3x_status q3x_copy_to_device(q3x_context ctx, void* dst, const void* src, unsigned long bytes);
q3x_status q3x_copy_to_host(q3x_context ctx, void* dst, const void* src, unsigned long bytes);

/* ---- per-kernel timing on the context stream (HIP events around each launch; the
 *      reference's ProfilingCounters / StageTimer, src/layer_0_core/profiling/counters.hpp).
 *      Kernel ids: 0 zsticks_inverse, 1 plane_sandwich, 2 zsticks_forward, 3 kb_overlap_gemm,
 *      4 kb_dblock, 5 kb_update_gemm; subspace eigensolver stages: 6 eig_prepare, 7 eig_gemms,
 *      8 eig_tridiagonal, 9 eig_bisection, 10 eig_inverse_iteration, 11 eig_back_transform. */
q3x_status q3x_profile_enable(q3x_context ctx, int enable);
/* synchronizes the stream; fills up to n totals (ms) and launch counts; reset != 0 zeroes them */
q3x_status q3x_profile_read(q3x_context ctx, int n, double* total_ms, long* count, int reset);
const char* q3x_profile_kernel_name(int id);

/* ---- fp64 generalized Hermitian eigensolver on host arrays (the Davidson subspace solve;
 *      stands in for GPUEigensolver::generalized_hermitian_eigenvalue_solve,
 *      src/layer_0_core/math/linear_algebra/gpu/lapack/init.hpp, with zhegv itype-1 semantics of
 *      subspace/serial.hpp:26-136): H x = w S x, upper triangles of H and S (n x n, column-major)
 *      read; all n eigenvalues ascending; the lowest nvec eigenvectors into X (n x nvec),
 *      X^H S X = I.  Returns Q3X_OK or Q3X_ERR_NUMERIC (S not positive definite). */
q3x_status q3x_hegv(int n, const double* H, const double* S, double* w, double* X, int nvec);
/* the same problem on device arrays (the GPU path of generalized_hermitian_eigenvalue_solve,
 * math/linear_algebra/gpu/lapack/init.hpp:77-82), solved on the device: H, S complex128 with
 * leading dimensions ldh, lds (upper triangles read); the lowest nvec eigenvalues ascending into
 * host w, their S-orthonormal eigenvectors into device X (n x nvec, ld ldx).  Synchronizes the
 * context stream; with q3x_profile_enable on, its stages are timed under kernel ids 6-11.
 * Returns Q3X_OK, Q3X_ERR_NUMERIC (S not positive definite) or Q3X_ERR_DEVICE. */
q3x_status q3x_hegv_device(q3x_context ctx, int n, const void* H, long ldh, const void* S, long lds,
                           int nvec, double* w, void* X, long ldx);
/* No counterpart in the reference.  Test hook of the device eigensolver: the next `count`
 * multi-workgroup tridiagonalisations (n >= 100) start with their grid-barrier give-up flag set,
 * so each runs to its end without waiting and its solve is redone on one workgroup, as when
 * another stream's work keeps a workgroup from being placed. */
void q3x_debug_force_barrier_giveups(int count);
/* No counterpart in the reference.  Test hook of the device eigensolver: bit i of mask set makes
 * the i-th eigensolve from now (q3x_hegv_device or one inside a Davidson solve, counted in the
 * process) report a non-positive-definite S (zhegv info > n) without solving, so that Davidson's
 * recovery paths (forced collapse, previous eigenvalues kept) can be run on purpose. */
void q3x_debug_fail_eigensolves(unsigned long long mask);
/* Solves redone on one workgroup since the process started. */
long q3x_eig_barrier_redos_total(void);

/* ---- block Davidson (PlanewaveDavidsonDiagonalizer::solve,
 *      src/layer_4_computation/diagonalizer/planewave/davidson.hpp:324-841, with the subspace
 *      manager of .../internal/subspace_manager/serial.hpp) */
typedef struct {
    double tolerance_ry;      /* set_tolerance; eigenvalue-change test (davidson.hpp:607-629) */
    int max_iterations;       /* set_maximum_iterations (default 20, davidson.hpp:1156) */
    int basis_factor;         /* set_basis_expansion_factor (default 2, davidson.hpp:864) */
    int n_occupied;           /* number_of_occupied_bands_per_spin */
    int precision;            /* Q3X_FP64 or Q3X_MIXED */
    int seed_from_psi;        /* 1: psi_inout seeds the basis (davidson.hpp:1000-1045);
                                 0: kinetic-sorted unit vectors (davidson.hpp:1047-1101) */
    int final_fp64_sweep;     /* MIXED only: finish with fp64 Rayleigh-Ritz + residual sweep */
    int stored_eigenvalues;   /* 1: eig_out holds the k-point state's stored eigenvalues on entry;
                                 when any is nonzero they are the previous eigenvalues of the
                                 first convergence test, else the diagonal of the starting H_sub
                                 is (davidson.hpp:473-493).  0: always the diagonal. */
} q3x_davidson_options;

typedef struct {
    int iterations;
    int unconverged;          /* number_of_unconverged_bands (davidson.hpp:270-276) */
    int h_applications;
    int fp64_refinements;     /* fp64 sweeps run after a mixed solve */
    double max_residual_norm; /* max_n ||H psi_n - e_n psi_n|| after the solve (fp64) */
    /* stage times in ms, named like DavidsonTimingAccumulator
     * (src/layer_0_core/profiling/counters.hpp:36-52):
     * 0 seed, 1 orthogonalize, 2 hamiltonian_application, 3 subspace_update,
     * 4 diagonalize, 5 collapse, 6 residual, 7 total.  Device times when profiling is enabled
     * on the context (q3x_profile_enable: each stage then ends with a stream synchronisation);
     * otherwise the solve synchronises only where the host needs device data and the stage
     * times are host-side. */
    double stage_ms[8];
    /* subspace eigensolves redone on one workgroup because a grid barrier of the
     * multi-workgroup tridiagonalisation gave up (other streams' work kept one of its
     * workgroups from being placed) */
    int eig_barrier_redos;
    /* recovery paths taken (davidson.hpp:520-590, subspace_manager/serial.hpp:370-552):
     * Cholesky-QR failures finished by modified Gram-Schmidt, subspace collapses forced by a
     * failed subspace eigensolve, and iterations that kept the previous eigenvalues because the
     * eigensolve failed again after such a collapse */
    int mgs_fallbacks;
    int recovery_collapses;
    int previous_eigenvalue_keeps;
} q3x_davidson_report;

void q3x_davidson_default_options(q3x_davidson_options* opt);
/* psi_inout: host, npw x nb complex; eig_out: nb (Ry, ascending). */
q3x_status q3x_davidson_solve(q3x_hamiltonian h, int nb, const q3x_davidson_options* opt,
                              double* psi_inout, double* eig_out, q3x_davidson_report* report);
/* the same with psi in HBM: d_psi_inout is a device buffer (npw x nb complex128, ld npw) on the
 * context's device; it seeds the solve when opt->seed_from_psi and receives the eigenvectors. */
q3x_status q3x_davidson_solve_device(q3x_hamiltonian h, int nb, const q3x_davidson_options* opt,
                                     void* d_psi_inout, double* eig_out, q3x_davidson_report* report);

/* ---- many k-points on one GPU: replaces the SCF k-loop of
 *      src/layer_4_computation/iterator/kohn_sham/planewave/pseudopotential/init.hpp:1056-1082
 *      (and the round-robin KPointDistributor, .../planewave/internal/kpoint_distributor.hpp)
 *      for independent k-points: k-point i runs on worker i % workers, each worker with its own
 *      context (stream, workspace) and host thread, so the workers' solves overlap. */
typedef struct q3x_batch_s* q3x_batch;
q3x_status q3x_batch_create(int device, int workers, q3x_batch* out);
q3x_status q3x_batch_destroy(q3x_batch b);
/* one k-point: a Hamiltonian with V(r) (as q3x_hamiltonian_set_potential), prepared for kp,
 * with its projector set (as q3x_hamiltonian_set_projectors; nproj may be 0).  *index = its
 * position in the batch. */
q3x_status q3x_batch_add(q3x_batch b, q3x_basis basis, q3x_kpoint kp, const double* v_r, long n,
                         const double* beta, int nproj, int natoms, const int* block_sizes,
                         const double* d_blocks, int* index);
int q3x_batch_size(q3x_batch b);
/* solves every k-point with the same options: psi_inout[i] host npw_i x nb, eig_out nk x nb
 * (row i = k-point i), reports (nk entries) may be null.  Returns the first failure. */
q3x_status q3x_batch_solve(q3x_batch b, int nb, const q3x_davidson_options* opt,
                           double* const* psi_inout, double* eig_out, q3x_davidson_report* reports);

/* ---- footprint estimate (PlanewaveMemoryEstimate::compute,
 *      src/layer_4_computation/iterator/kohn_sham/planewave/internal/memory_estimate.hpp:37-215):
 *      the reference's inputs, terms (label, formula, bytes), integer arithmetic and totals.
 *      Host-only. */
typedef struct {
    int number_of_bands;
    int max_plane_waves_per_kpoint;
    int number_of_g_vectors;
    int number_of_projectors;
    int number_of_atomic_orbitals;
    int number_of_atoms;
    int number_of_spin_components;
    int local_number_of_kpoint_states;
    int mixing_history_depth;
    long long fft_rho_grid_points;
    long long fft_wave_grid_points;
    int davidson_basis_expansion_factor;
    int overlap_is_identity;       /* false for ultrasoft and PAW (S stored) */
    int diagonalizer_is_davidson;  /* false for the direct solver */
} q3x_memory_estimate_inputs;
typedef struct {
    const char* label;    /* static strings */
    const char* formula;
    unsigned long long bytes;
} q3x_memory_term;
typedef struct {
    q3x_memory_term steady_state[16];
    int n_steady_state;
    q3x_memory_term transient[4];
    int n_transient;
    unsigned long long steady_state_bytes, solver_bytes, setup_peak_bytes, peak_bytes;
    int maximum_basis_size;
} q3x_memory_estimate;
/* the reference's defaults (PlanewaveMemoryEstimateInputs{}): zeros, one spin, one k-point
 * state, expansion factor 2, S not the identity, the Davidson solver */
void q3x_memory_estimate_default_inputs(q3x_memory_estimate_inputs* in);
q3x_status q3x_memory_estimate_compute(const q3x_memory_estimate_inputs* in, q3x_memory_estimate* out);

/* ---- band-structure request server (continuous batching).  No counterpart in the reference,
 *      which runs one calculation per process; it serves non-self-consistent band solves of the
 *      reference's kind (flavors/non_self_consistent_field.hpp:78-112) for many cells at once.
 *      Requests are admitted in arrival order while their estimated HBM footprint (the terms of
 *      PlanewaveMemoryEstimate::compute, iterator/kohn_sham/planewave/internal/
 *      memory_estimate.hpp:75-215) fits the budget; the k-points of admitted requests share the
 *      workers (one context, stream and host thread each), so requests that arrive while others
 *      run overlap with them on the device.  Wavefunctions stay in HBM until the result is
 *      collected. */
typedef struct q3x_server_s* q3x_server;
typedef struct {
    q3x_basis basis;          /* cell and cutoff; caller-owned, alive until q3x_server_wait returns */
    const double* v_r;        /* V(r) on the basis' wave grid, as q3x_hamiltonian_set_potential */
    long nv;
    int nk;                   /* k-points, crystal coordinates, nk x 3 */
    const double* kpoints;
    int natoms;               /* atoms and species, as q3x_kb_projectors */
    const double* atom_pos;
    const int* atom_species;
    int nspecies;
    const int* nmesh;
    const double* r_all;
    const double* rab_all;
    const int* nbeta;
    const int* l_all;
    const int* cutoff_all;
    const double* beta_r_all;
    const int* block_sizes;   /* D blocks per atom, as q3x_hamiltonian_set_projectors */
    const double* d_blocks;
    int nbands;
    int noccupied;
    q3x_davidson_options options;  /* tolerance, iterations, precision (seed_from_psi and
                                      stored_eigenvalues are set by the server) */
    /* The k-loop of the reference's non-self-consistent run over a band path
     * (iterator/kohn_sham/planewave/pseudopotential/init.hpp:1036-1106, 1599-1815).
     * kpath_propagation 0: the k-points are independent, each starts from
     *   q3x_hamiltonian_random_psi with its index as the seed, and they run concurrently on the
     *   server's workers.
     * kpath_propagation 1: the k-points are solved in order on one worker; k-point k starts
     *   from k-1's eigenvectors (the leading min(npw) coefficients of each band, zeros after them)
     *   and eigenvalues unless k-1 is a band-path segment boundary (a multiple of
     *   band_path_points_per_segment below band_path_segments x points; none without a band
     *   path), else from its own random start; then, when nk > 2, the backward fix re-solves
     *   k-points edge+5 .. edge in descending order for edge 0 and the first k-point of every
     *   later segment, each seeded by the k-point after it. */
    int kpath_propagation;
    int band_path_segments;            /* band_path_.size() - 1; 0: no band path */
    int band_path_points_per_segment;
} q3x_band_request;
typedef struct {
    double submit_s, admit_s, done_s;  /* seconds since the server started */
    double footprint_bytes;            /* the admission estimate */
    int iterations;                    /* Davidson iterations summed over the k-points */
    double setup_s;                    /* worker time in per-k-point setup (G list, projectors, upload) */
    double solve_s;                    /* worker time in the Davidson solves */
    int eig_barrier_redos;             /* subspace eigensolves redone on one workgroup (see
                                          q3x_davidson_report), summed over the k-points */
    int kpoints_done;                  /* k-points solved (fewer than requested after a cancel) */
    int solves;                        /* Davidson solves run, the backward fix's re-solves included */
} q3x_request_stats;
/* hbm_budget_bytes <= 0: 80% of the device memory free at creation */
q3x_status q3x_server_create(int device, int workers, double hbm_budget_bytes, q3x_server* out);
/* waits for every submitted request, then stops the workers */
q3x_status q3x_server_destroy(q3x_server s);
/* Drops every k-point not yet started: requests not yet admitted, tasks of admitted requests not
 * yet handed out, prepared k-points not yet taken by a worker.  k-points already on the device
 * finish.  Each affected request completes with Q3X_ERR_STATE ("cancelled ...") from
 * q3x_server_wait once its running k-points are done; *cancelled (may be null) counts them.  No
 * counterpart in the reference (it serves one calculation per process). */
q3x_status q3x_server_cancel_pending(q3x_server s, int* cancelled);
q3x_status q3x_server_estimate(q3x_server s, const q3x_band_request* req, double* bytes);
/* The admission footprint of a request, itemised; host-only (needs no server or device).
 * reference_out: PlanewaveMemoryEstimate::compute of the request's dimensions (bands, the first
 * k-point's plane waves plus 5% and 16, G vectors, projectors, atoms, nk wavefunction sets, the
 * boxes, the expansion factor; no atomic orbitals, no mixing history, S = identity).  The
 * footprint counts the reference terms this path holds in HBM -- "wavefunctions" once, and
 * "Davidson basis V, HV", "rotation scratch", "subspace matrices", "nonlocal projectors" and
 * "block-diagonal D matrix" per concurrently running k-point -- plus device_terms, the buffers
 * the reference's estimate has no line for, also per concurrently running k-point; the
 * reference's host-side lines (G-vector table, FFT buffers, density and potential, Broyden
 * history) and its atomic-orbital start are not on the device.
 * concurrent_kpoints: 1 for a k-path request, else the server's worker count (the estimate
 * q3x_server_estimate uses).  device_terms holds at most 8 entries. */
q3x_status q3x_band_request_footprint(const q3x_band_request* req, int concurrent_kpoints,
                                      q3x_memory_estimate* reference_out, q3x_memory_term* device_terms,
                                      int* n_device_terms, double* bytes);
/* copies the request's arrays and returns at once with a ticket */
q3x_status q3x_server_submit(q3x_server s, const q3x_band_request* req, long* ticket);
/* blocks until the request is done: eig_out nk x nb (Ry, row k = k-point k); frees its state */
q3x_status q3x_server_wait(q3x_server s, long ticket, double* eig_out, q3x_request_stats* stats);
double q3x_server_in_flight_bytes(q3x_server s);

#ifdef __cplusplus
}
#endif
#endif /* Q3X_H */
