// Device runtime objects behind the C ABI: context, device buffers, Hamiltonian state.
#pragma once
#include <hip/hip_runtime.h>

#include <memory>
#include <string>
#include <vector>

#include "fft_lds.h"
#include "hlocal.h"
#include "q3x_internal.h"

namespace q3x {

bool hip_ok(hipError_t e, const char* what);

// Owning device allocation; grows, never shrinks.
class DevBuf {
  public:
    DevBuf() = default;
    DevBuf(const DevBuf&) = delete;
    DevBuf& operator=(const DevBuf&) = delete;
    ~DevBuf() { release(); }
    bool ensure(size_t bytes);
    void release();
    template <typename T>
    T* as() const {
        return static_cast<T*>(p_);
    }
    void* get() const { return p_; }
    size_t size() const { return bytes_; }
    // Copies bytes from host memory in, growing the buffer.  With pin (and up to 16 MiB) the
    // copy goes through that pinned staging and returns without waiting; otherwise it waits.
    bool upload(const void* src, size_t bytes, hipStream_t st, class HostStage* pin = nullptr);

  private:
    void* p_ = nullptr;
    size_t bytes_ = 0;
};

// Pinned host memory for the small copies of a solve (eigenvalues, Gram blocks, index lists).
// A copy between device memory and pageable host memory goes through the runtime's shared
// staging buffer and holds it until its stream reaches the copy, which serialises the solves of
// other host threads; from pinned memory it is a plain DMA on the stream.  Regions are handed
// out as a ring: take() synchronises the stream and rewinds when the ring is full (so no region
// is reused while a copy from it may be pending), and d2h() rewinds after its synchronisation.
class HostStage {
  public:
    HostStage() = default;
    HostStage(const HostStage&) = delete;
    HostStage& operator=(const HostStage&) = delete;
    ~HostStage();
    char* take(size_t bytes, hipStream_t st);  // nullptr on failure
    void rewind() { off_ = 0; }                 // only right after a synchronisation of the stream
    hipError_t h2d(void* dst, const void* src, size_t bytes, hipStream_t st);  // asynchronous
    hipError_t d2h(void* dst, const void* src, size_t bytes, hipStream_t st);  // waits for the stream

  private:
    char* p_ = nullptr;
    size_t cap_ = 0, off_ = 0;
};

// Per-kernel-family timing with HIP events recorded on the context stream around each launch.
enum KernelId {
    K_ZINV = 0, K_PLANE, K_ZFWD, K_KB_OVERLAP, K_KB_DBLOCK, K_KB_UPDATE,
    // stages of the subspace eigensolver
    K_EIG_PREP, K_EIG_GEMM, K_EIG_TRIDIAG, K_EIG_BISECT, K_EIG_INVIT, K_EIG_BACK, K_COUNT
};
struct Profiler {
    bool on = false;
    std::vector<hipEvent_t> pool;
    size_t used = 0;
    struct Span {
        int id, e0, e1;
    };
    std::vector<Span> spans;
    double total_ms[K_COUNT] = {};
    long count[K_COUNT] = {};
    int mark(hipStream_t st);  // records the next pool event, returns its index (-1 if off)
    void span(int id, int e0, int e1) {
        if (on && e0 >= 0 && e1 >= 0) spans.push_back({id, e0, e1});
    }
    bool flush(hipStream_t st);
    ~Profiler();
};

struct FftTables {
    FftPlan plan;
    DevBuf tw64, tw32;        // stage-major Stockham twiddles
    DevBuf plain64, plain32;  // exp(-2 pi i j/n), j < n
};

// Device buffers of the block Davidson (the reference's SerialDavidsonWorkspace,
// davidson.hpp:120-221), kept by the context and grown on demand, so that repeated solves on a
// context allocate nothing once the largest size has been seen.
struct DavidsonWorkspace {
    // V2 / HV2: the other half of the basis ping-pong (a rotation writes V X there and swaps)
    DevBuf V, HV, V2, HV2, R, C, G, Hs, Ss, X, Xc, eps, negeps, work, norms, idx, eigw;
};

}  // namespace q3x

struct q3x_context_s {
    q3x::Profiler prof;
    q3x::DevBuf gemm_work;  // split-K workspace of q3x_zgemm
    q3x::DevBuf eig_work;   // workspace of q3x_hegv_device
    q3x::DavidsonWorkspace dav;
    q3x::HostStage pin;  // pinned staging of the small host copies
    int device = 0;
    hipStream_t stream = nullptr;
    size_t max_lds = 65536;
    int num_cus = 0;
};

struct q3x_hamiltonian_s {
    q3x_context ctx = nullptr;
    const q3x_basis_s* basis = nullptr;
    int dims[3] = {0, 0, 0};
    long npoints = 0;
    q3x::FftTables fft[3];
    // potential V(r)/N, plane-major [i3][i2][i1]
    q3x::DevBuf v64, v32;
    bool have_potential = false;
    // prepared k-point
    const q3x_kpoint_s* kp = nullptr;
    int npw = 0, nst = 0, ncol = 0, n_outside = 0;
    int sticks_per_block = 16, rows_per_pass64 = 8, rows_per_pass32 = 8;
    q3x::PlaneTlConfig tl64, tl32;  // two-level plane kernel (ok = false: generic kernel)
    q3x::StickTlConfig tz64, tz32;  // two-level z-stick kernels (ok = false: generic kernels)
    q3x::DevBuf stick_off, stick_i2, stick_col, col_i1, ent_pw, ent_pos, ekin, outside;
    // projectors of the prepared k-point
    int nproj = 0;
    bool have_projectors = false;
    q3x::DevBuf beta64, beta32, dblocks, pstart, psize, poff;
    // inputs of the device projector builder (q3x_hamiltonian_build_projectors): G list of the
    // k-point, beta(q) tables, form factors, channel / atom tables
    q3x::DevBuf kb_g, kb_tables, kb_form, kb_ints, kb_dbl;
    // Gamma trick: transforms run over the extended list (half sphere, then mirrors) of npw_ext
    // entries, two bands per column: gpack holds the packed pairs, gout their H images
    bool gamma = false;
    int npw_ext = 0;
    q3x::DevBuf gpack, gout;
    q3x::DevBuf dens_w;  // per-column weights of q3x_hamiltonian_accumulate_density
    // scratch
    q3x::DevBuf zbuf, P, W, gemm_work, io_psi, io_hpsi;
    q3x::DevBuf plane_slots;  // per-workgroup planes of the two-level kernel when a plane exceeds LDS
};

namespace q3x {
// Points an existing Hamiltonian at another basis (tables rebuilt, k-point, potential and
// projectors cleared) keeping its device buffers, which only grow: a long-lived worker reuses one
// Hamiltonian across requests without allocating or freeing device memory once warm.
q3x_status hamiltonian_rebind(q3x_hamiltonian h, q3x_basis b);
// Device bytes a context (its workspaces) and a Hamiltonian (may be null) hold.
double device_bytes_held(const q3x_context_s* ctx, const q3x_hamiltonian_s* h);
// Applies H to nvec device vectors (ld = npw) on the context stream.
q3x_status hamiltonian_apply_device(q3x_hamiltonian h, const double* d_psi, long ldpsi, double* d_hpsi, long ldh,
                                    int nvec, int precision);
// The same product with every scratch buffer (stick buffer, plane slots, projector products,
// split-K workspace) carved from one caller-given device region: bands run in chunks that fit
// it; nothing is allocated.  Fails with Q3X_ERR_ARGUMENT when one band does not fit.
q3x_status hamiltonian_apply_fixed(q3x_hamiltonian h, const double* d_psi, double* d_hpsi, int nvec, int precision,
                                   char* scratch, size_t scratch_bytes);
// D blocks of the projector set of the prepared k-point (packed per atom, block_sizes summing to
// nproj); marks the set complete (beta64 / beta32 must already hold the projectors).
q3x_status upload_dblocks(q3x_hamiltonian h, int nproj, int natoms, const int* block_sizes, const double* d_blocks);
// Looks up, once per device and with that device current, the device address of the fixed
// scratch region of q3x_hamiltonian_apply_noalloc (noalloc.hip).
bool noalloc_scratch_init(int device);
}  // namespace q3x
