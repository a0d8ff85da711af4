// Band-structure request server: continuous batching of non-self-consistent solves on one GPU.
//
// A request is one cell (basis, V(r), atoms and projector data) with a list of k-points; its
// result is nk x nb eigenvalues, its wavefunctions stay in HBM until the client collects the
// result.  Requests are admitted in arrival order while their estimated HBM footprint fits the
// budget (the reference's PlanewaveMemoryEstimate::compute terms,
// iterator/kohn_sham/planewave/internal/memory_estimate.hpp:75-215); an admitted request's
// k-points join one shared task queue.  P preparation threads take k-point tasks from it in
// order and do their host work (the k+G list, the box maps and the stick layout); W worker
// threads -- each with its own context, stream and Davidson workspace -- take prepared k-points,
// form the KB projector set and the seeded starting block on the device and solve, so the host
// setup of the next k-points overlaps the device work of the current ones.  A request that arrives
// while others are running is admitted as soon as the budget allows and its k-points interleave
// with theirs on the device.  What admission counts as resident: the wavefunctions of every
// admitted request not yet collected, the blocks parked in the reuse pool, and the larger of the
// running requests' per-k-point shares and the workers' own buffers (Davidson workspace and
// Hamiltonian, which persist between k-points).  A request's per-k-point share is released when
// its last k-point finishes, its wavefunction share when the client collects the result; when
// the head of the queue does not fit, the pool is emptied first.
//
// A k-path request (kpath_propagation 1) follows the reference's non-self-consistent k-loop
// (iterator/kohn_sham/planewave/pseudopotential/init.hpp:1036-1106): each k-point starts from its
// predecessor's converged wavefunctions and eigenvalues, and the edge k-points are re-solved
// backwards after the forward pass.  Its solves depend on each other, so the request is one task:
// one worker runs its k-loop in order while the other workers serve other requests.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <numeric>
#include <stdexcept>
#include <string>
#include <thread>
#include <vector>

#include "runtime.h"

namespace {

using clk = std::chrono::steady_clock;

struct Request {
    long ticket = 0;
    q3x_basis basis = nullptr;
    std::vector<double> v, kpts, atom_pos, r_all, rab_all, beta_r_all, d_blocks;
    std::vector<int> atom_species, nmesh, nbeta, l_all, cutoff_all, block_sizes;
    int nk = 0, natoms = 0, nspecies = 0, nb = 0, nocc = 0;
    q3x_davidson_options opt{};
    double bytes = 0.0;      // admission estimate: bytes_psi + per-k-point share
    double bytes_psi = 0.0;  // wavefunctions of all k-points, held until the result is collected
    // progress
    int remaining = 0;
    int next_task = 0;
    bool admitted = false, done = false;
    q3x_status status = Q3X_OK;
    std::string error;
    std::vector<double> eig;
    std::vector<std::pair<q3x_context, void*>> psi;  // device wavefunctions, one per k-point
    std::vector<size_t> psi_bytes;
    int iterations = 0;
    int eig_barrier_redos = 0;
    int kpoints_done = 0;
    int solves = 0;
    // k-path requests (kpath_propagation 1): the k-loop's solves in order, (target, seed or -1
    // for the k-point's own random start); one worker runs them, so the request is one task
    bool kpath = false;
    std::vector<std::pair<int, int>> steps;
    bool cancel = false;  // set by cancel_pending: the k-path stops before its next solve
    double t_submit = 0, t_admit = 0, t_done = 0;
    double setup_s = 0, solve_s = 0;
};

double now_s(clk::time_point t0) { return std::chrono::duration<double>(clk::now() - t0).count(); }

// is_at_band_path_segment_boundary (iterator/kohn_sham/planewave/pseudopotential/init.hpp:1697-1715)
bool segment_boundary(int k, int segments, int pps) {
    if (segments <= 0 || pps <= 0 || k >= segments * pps) return false;
    return k % pps == 0;
}

// The solves of the non-self-consistent k-loop in order (pseudopotential/init.hpp:1036-1106):
// the forward pass, k seeded by k-1 unless k-1 is a segment boundary
// (propagate_wavefunctions_to_next_kpoint, :1599-1645), then for nk > 2 the backward fix
// (fix_edge_kpoints_with_backward_propagation, :1735-1800) over edge 0 and the first k-point of
// every later segment: edge+5 .. edge in descending order, the first seeded by the k-point after
// the sweep (or the last k-point), each next one by the one just solved.
std::vector<std::pair<int, int>> nscf_schedule(int nk, int segments, int pps) {
    std::vector<std::pair<int, int>> steps;
    for (int k = 0; k < nk; ++k)
        steps.emplace_back(k, k > 0 && !segment_boundary(k - 1, segments, pps) ? k - 1 : -1);
    if (nk <= 2) return steps;
    std::vector<int> edges{0};
    if (segments > 0 && pps > 0)
        for (int sg = 1; sg < segments; ++sg)
            if (sg * pps < nk) edges.push_back(sg * pps);
    constexpr int kSweepDepth = 5;  // backward_sweep_depth, :1767
    for (int e : edges) {
        const int end = std::min(e + kSweepDepth, nk - 1);
        int seed = std::min(end + 1, nk - 1);
        for (int t = end; t >= e; --t) {
            steps.emplace_back(t, seed);
            seed = t;
        }
    }
    return steps;
}

// The host part of a k-point task, done ahead of the device part by a preparation thread.
struct Prepared {
    std::shared_ptr<Request> r;
    int k = -1;
    q3x_kpoint kp = nullptr;
    int npw = 0;
    q3x_status status = Q3X_OK;
    std::string err;
    double host_s = 0.0;
};

}  // namespace

struct q3x_server_s {
    int device = 0;
    double budget = 0.0;
    // resident-memory terms of admission (see the file comment)
    double psi_held = 0.0, work_running = 0.0, pool_bytes = 0.0;
    std::vector<double> worker_bytes;  // device bytes each worker's context and Hamiltonian hold
    double workers_locked() const {
        double ws = 0.0;
        for (double b : worker_bytes) ws += b;
        return ws;
    }
    double resident_locked() const { return psi_held + pool_bytes + std::max(workers_locked(), work_running); }
    // resident memory once r is admitted: r's per-k-point work may reuse the workers' buffers
    double resident_with_locked(const Request& r) const {
        return psi_held + pool_bytes + r.bytes_psi + std::max(workers_locked(), work_running + (r.bytes - r.bytes_psi));
    }
    clk::time_point t0;
    std::vector<q3x_context> workers;
    std::vector<std::thread> threads, prep_threads;
    std::mutex mu;
    std::condition_variable cv_task, cv_done, cv_ready;
    std::deque<std::unique_ptr<Prepared>> ready;  // prepared k-points waiting for a worker
    int preparing = 0;                           // tasks taken by preparation threads, not yet ready
    size_t max_ready = 16;                       // bound on prepared-but-unsolved k-points
    std::deque<std::shared_ptr<Request>> waiting;   // submitted, not admitted
    std::deque<std::shared_ptr<Request>> running;   // admitted, tasks left to hand out
    std::map<long, std::shared_ptr<Request>> all;
    long next_ticket = 1;
    bool stop = false;
    // one Hamiltonian per worker, rebound to each k-point's cell; device buffers only grow
    std::vector<q3x_hamiltonian> hams;
    std::vector<long> ham_ticket;  // request whose cell and potential hams[w] holds (0: none)
    // device wavefunction blocks returned by finished requests, by size, for reuse: after warm-up
    // a k-point takes a block from here instead of allocating (hipFree synchronises the device)
    std::multimap<size_t, void*> psi_pool;

    // a block of at least `bytes`; *held = its size
    void* take_psi(q3x_context ctx, size_t bytes, size_t* held) {
        {
            std::lock_guard<std::mutex> g(mu);
            auto it = psi_pool.lower_bound(bytes);
            if (it != psi_pool.end()) {
                void* p = it->second;
                *held = it->first;
                pool_bytes -= static_cast<double>(it->first);
                psi_pool.erase(it);
                return p;
            }
        }
        void* p = nullptr;
        *held = bytes;
        return q3x_device_alloc(ctx, static_cast<unsigned long>(bytes), &p) == Q3X_OK ? p : nullptr;
    }

    // admission in arrival order: the head waits until it fits, or until no other request holds
    // memory; parked wavefunction blocks are freed before the head is made to wait for them
    void admit_locked() {
        while (!waiting.empty()) {
            auto r = waiting.front();
            if (resident_with_locked(*r) > budget && !psi_pool.empty()) {
                for (auto& kv : psi_pool) (void)hipFree(kv.second);
                psi_pool.clear();
                pool_bytes = 0.0;
            }
            const bool others = psi_held > 0.0 || work_running > 0.0;
            if (others && resident_with_locked(*r) > budget) break;
            waiting.pop_front();
            r->admitted = true;
            r->t_admit = now_s(t0);
            psi_held += r->bytes_psi;
            work_running += r->bytes - r->bytes_psi;
            running.push_back(r);
        }
        cv_task.notify_all();
    }

    // preparation thread: next task in order -> host work -> ready queue
    void preparer() {
        for (;;) {
            std::shared_ptr<Request> r;
            int k = -1;
            {
                std::unique_lock<std::mutex> g(mu);
                cv_task.wait(g, [&] {
                    return stop || (!running.empty() && ready.size() + static_cast<size_t>(preparing) < max_ready);
                });
                if (stop) return;
                r = running.front();
                if (r->kpath) {  // one task; the worker prepares its k-points itself
                    r->next_task = r->nk;
                } else {
                    k = r->next_task++;
                }
                if (r->next_task >= r->nk) running.pop_front();
                ++preparing;
            }
            std::unique_ptr<Prepared> p;
            try {
                p = prepare(r, k);
            } catch (const std::exception& ex) {  // e.g. an allocation: the k-point fails, the thread lives
                p = std::make_unique<Prepared>();
                p->r = r;
                p->k = k;
                p->status = Q3X_ERR_INTERNAL;
                p->err = std::string("preparing the k-point: ") + ex.what();
            }
            std::lock_guard<std::mutex> g(mu);
            --preparing;
            ready.push_back(std::move(p));
            cv_ready.notify_one();
        }
    }

    std::unique_ptr<Prepared> prepare(const std::shared_ptr<Request>& r, int k) {
        const clk::time_point ts = clk::now();
        auto p = std::make_unique<Prepared>();
        p->r = r;
        p->k = k;
        if (k < 0) return p;  // a k-path request
        auto fail = [&](q3x_status st) {
            p->status = st;
            p->err = q3x_last_error();
            return std::move(p);
        };
        // the host part of a k-point is its plane-wave list, box maps and stick layout; its
        // projector set and starting block are formed on the device by the worker
        q3x_status s = q3x_kpoint_create(r->basis, &r->kpts[3 * k], &p->kp, &p->npw);
        if (s != Q3X_OK) return fail(s);
        p->host_s = std::chrono::duration<double>(clk::now() - ts).count();
        return p;
    }

    void worker(int w) {
        q3x_context ctx = workers[w];
        // a worker that cannot select its device still takes k-points, and fails them, so that
        // their requests finish and waiters are released
        const bool device_ok = hipSetDevice(device) == hipSuccess;
        for (;;) {
            std::unique_ptr<Prepared> p;
            {
                std::unique_lock<std::mutex> g(mu);
                cv_ready.wait(g, [&] { return stop || !ready.empty(); });
                if (ready.empty()) return;  // stop, and nothing prepared is left
                p = std::move(ready.front());
                ready.pop_front();
                cv_task.notify_all();  // room in the ready queue
            }
            Request& r = *p->r;
            std::string err;
            q3x_status s = p->status;
            if (s != Q3X_OK) {
                err = p->err;
            } else if (!device_ok) {
                s = Q3X_ERR_DEVICE;
                err = "the worker could not select device " + std::to_string(device);
            } else {
                s = p->k < 0 ? solve_path(w, ctx, r, err) : solve_kpoint(w, ctx, *p, err);
            }
            if (p->kp) q3x_kpoint_destroy(p->kp);
            const double held = device_ok ? q3x::device_bytes_held(ctx, hams[w]) : 0.0;
            std::lock_guard<std::mutex> g(mu);
            worker_bytes[w] = held;
            if (s != Q3X_OK && r.status == Q3X_OK) {
                r.status = s;
                r.error = (p->k >= 0 ? "k-point " + std::to_string(p->k) + ": " : std::string()) + err;
            }
            if (--r.remaining == 0) {
                r.done = true;
                r.t_done = now_s(t0);
                work_running -= r.bytes - r.bytes_psi;
                if (work_running < 1.0) work_running = 0.0;
                admit_locked();
                cv_done.notify_all();
            }
        }
    }

    // bind worker w's Hamiltonian to r's cell (kept while the worker's k-points come from the same
    // request) and to kp, with kp's projector set formed on the device (form factors x structure
    // factors; the beta(q) tables of the cell come from the shared cache)
    q3x_status bind(int w, q3x_context ctx, Request& r, q3x_kpoint kp) {
        q3x_hamiltonian& h = hams[w];
        q3x_status s = Q3X_OK;
        if (!(h && ham_ticket[w] == r.ticket)) {
            ham_ticket[w] = 0;
            if ((s = h ? q3x::hamiltonian_rebind(h, r.basis) : q3x_hamiltonian_create(ctx, r.basis, &h)) != Q3X_OK ||
                (s = q3x_hamiltonian_set_potential(h, r.v.data(), static_cast<long>(r.v.size()))) != Q3X_OK)
                return s;
            ham_ticket[w] = r.ticket;
        }
        if ((s = q3x_hamiltonian_prepare_kpoint(h, kp)) != Q3X_OK) return s;
        return q3x_hamiltonian_build_projectors(h, r.natoms, r.atom_pos.data(), r.atom_species.data(), r.nspecies,
                                                r.nmesh.data(), r.r_all.data(), r.rab_all.data(), r.nbeta.data(),
                                                r.l_all.data(), r.cutoff_all.data(), r.beta_r_all.data(),
                                                r.block_sizes.data(), r.d_blocks.data());
    }

    // k-point k's device wavefunction block, taken on its first solve
    void* psi_block(q3x_context ctx, Request& r, int k, int npw) {
        {
            std::lock_guard<std::mutex> g(mu);
            if (r.psi[k].second) return r.psi[k].second;
        }
        size_t held = 0;
        void* psi = take_psi(ctx, static_cast<size_t>(npw) * r.nb * 16, &held);
        if (!psi) return nullptr;
        std::lock_guard<std::mutex> g(mu);
        r.psi[k] = {ctx, psi};
        r.psi_bytes[k] = held;
        return psi;
    }

    // Davidson on k-point k in place on its block (seeded by the block); eigenvalues into row k
    q3x_status davidson(int w, Request& r, int k, void* psi, bool stored, clk::time_point ts, double host_s,
                        bool first_solve) {
        q3x_davidson_options o = r.opt;
        o.seed_from_psi = 1;
        o.stored_eigenvalues = stored ? 1 : 0;
        q3x_davidson_report rep{};
        const clk::time_point tsolve = clk::now();
        const q3x_status s = q3x_davidson_solve_device(hams[w], r.nb, &o, psi, &r.eig[static_cast<size_t>(k) * r.nb], &rep);
        if (s != Q3X_OK) return s;
        const clk::time_point tend = clk::now();
        std::lock_guard<std::mutex> g(mu);
        r.iterations += rep.iterations;
        r.eig_barrier_redos += rep.eig_barrier_redos;
        r.kpoints_done += first_solve ? 1 : 0;
        ++r.solves;
        r.setup_s += host_s + std::chrono::duration<double>(tsolve - ts).count();
        r.solve_s += std::chrono::duration<double>(tend - tsolve).count();
        return Q3X_OK;
    }

    // device part of an independent k-point task: bind, random start, solve
    q3x_status solve_kpoint(int w, q3x_context ctx, Prepared& p, std::string& err) {
        const clk::time_point ts = clk::now();
        Request& r = *p.r;
        auto fail = [&](q3x_status st) {
            err = q3x_last_error();
            return st;
        };
        q3x_status s = bind(w, ctx, r, p.kp);
        if (s != Q3X_OK) return fail(s);
        void* psi = psi_block(ctx, r, p.k, p.npw);
        if (!psi) return fail(Q3X_ERR_DEVICE);
        // starting block: q3x_kpoint_random_psi with the k-point's index as the seed, formed on the
        // device, so a request's result does not depend on what else runs (plane-wave unit vectors,
        // the solver's own seed, can miss symmetry sectors of these undisplaced crystals)
        if ((s = q3x_hamiltonian_random_psi(hams[w], r.nb, static_cast<unsigned long long>(p.k), psi)) != Q3X_OK)
            return fail(s);
        if ((s = davidson(w, r, p.k, psi, false, ts, p.host_s, true)) != Q3X_OK) return fail(s);
        return Q3X_OK;
    }

    // a k-path request: its k-loop's solves in order on this worker (nscf_schedule).  A seeded
    // solve starts from the seed's eigenvectors, the leading min(npw) coefficients of each band
    // with zeros after them, and the seed's eigenvalues as the stored ones
    // (pseudopotential/init.hpp:1599-1688); an unseeded one from the k-point's random start.
    q3x_status solve_path(int w, q3x_context ctx, Request& r, std::string& err) {
        std::vector<q3x_kpoint> kps(r.nk, nullptr);
        std::vector<int> npw(r.nk, 0);
        std::vector<char> solved(r.nk, 0);
        q3x_status s = Q3X_OK;
        for (const auto& step : r.steps) {
            {
                std::lock_guard<std::mutex> g(mu);
                if (r.cancel) break;
            }
            const int t = step.first, seed = step.second;
            const clk::time_point ts = clk::now();
            double host_s = 0.0;
            if (!kps[t]) {
                if ((s = q3x_kpoint_create(r.basis, &r.kpts[3 * static_cast<size_t>(t)], &kps[t], &npw[t])) != Q3X_OK)
                    break;
                host_s = std::chrono::duration<double>(clk::now() - ts).count();
            }
            if ((s = bind(w, ctx, r, kps[t])) != Q3X_OK) break;
            void* psi = psi_block(ctx, r, t, npw[t]);
            if (!psi) {
                s = Q3X_ERR_DEVICE;
                break;
            }
            double* eig_t = &r.eig[static_cast<size_t>(t) * r.nb];
            if (seed < 0) {
                if ((s = q3x_hamiltonian_random_psi(hams[w], r.nb, static_cast<unsigned long long>(t), psi)) != Q3X_OK)
                    break;
            } else if (seed != t) {
                const size_t ns = static_cast<size_t>(npw[seed]), nt = static_cast<size_t>(npw[t]);
                const size_t nc = std::min(ns, nt);
                hipStream_t st = ctx->stream;
                if (!q3x::hip_ok(hipMemcpy2DAsync(psi, nt * 16, r.psi[seed].second, ns * 16, nc * 16, r.nb,
                                             hipMemcpyDeviceToDevice, st),
                            "propagating wavefunctions") ||
                    (nt > nc && !q3x::hip_ok(hipMemset2DAsync(static_cast<char*>(psi) + nc * 16, nt * 16, 0, (nt - nc) * 16,
                                                         r.nb, st),
                                        "zero-padding wavefunctions"))) {
                    s = Q3X_ERR_DEVICE;
                    break;
                }
                std::copy(&r.eig[static_cast<size_t>(seed) * r.nb], &r.eig[static_cast<size_t>(seed + 1) * r.nb], eig_t);
            }
            if ((s = davidson(w, r, t, psi, true, ts, host_s, !solved[t])) != Q3X_OK) break;
            solved[t] = 1;
        }
        if (s != Q3X_OK) err = q3x_last_error();
        for (q3x_kpoint kp : kps)
            if (kp) q3x_kpoint_destroy(kp);
        {
            std::lock_guard<std::mutex> g(mu);
            if (s == Q3X_OK && r.cancel) {
                s = Q3X_ERR_STATE;
                err = "cancelled before all k-points were solved";
            }
        }
        return s;
    }
};

using namespace q3x;

namespace {

template <typename T>
std::vector<T> copy_of(const T* p, size_t n) {
    return p ? std::vector<T>(p, p + n) : std::vector<T>(n, T{});
}

// Admission footprint of a request (q3x_band_request_footprint): PlanewaveMemoryEstimate::compute
// (memory_estimate.hpp:75-215) of the request's dimensions, of which the lines this path holds in
// HBM count -- the wavefunctions of every k-point once, the solver's lines, the projectors and the
// D blocks per concurrently running k-point -- plus the device buffers that estimate has no line
// for, also per running k-point.
struct Footprint {
    q3x_memory_estimate ref{};
    q3x_memory_term dev[8]{};
    int ndev = 0;
    double psi = 0.0, per_kpoint = 0.0, total = 0.0;
};

q3x_status request_footprint(const q3x_band_request* q, int concurrent, Footprint& f) {
    if (!q || !q->basis || !q->kpoints || q->nk < 1 || q->nbands < 1 || concurrent < 1 || q->natoms < 0 ||
        (q->natoms > 0 && !q->block_sizes)) {
        set_error("q3x_band_request_footprint: invalid argument");
        return Q3X_ERR_ARGUMENT;
    }
    q3x_kpoint kp = nullptr;
    int npw = 0;
    const q3x_status st = q3x_kpoint_create(q->basis, q->kpoints, &kp, &npw);
    if (st != Q3X_OK) return st;
    q3x_kpoint_destroy(kp);
    int rho[3], wave[3], ng = 0;
    double vol = 0, tpa2 = 0;
    q3x_basis_info(q->basis, rho, wave, &ng, &vol, &tpa2);
    int nproj = 0;
    for (int a = 0; a < q->natoms; ++a) nproj += q->block_sizes[a];
    // the first k-point's plane waves stand for the path's; the others may hold a few more
    const int npw_est = static_cast<int>(npw * 1.05) + 16;
    q3x_memory_estimate_inputs in;
    q3x_memory_estimate_default_inputs(&in);
    in.number_of_bands = q->nbands;
    in.max_plane_waves_per_kpoint = npw_est;
    in.number_of_g_vectors = ng;
    in.number_of_projectors = nproj;
    in.number_of_atoms = q->natoms;
    in.local_number_of_kpoint_states = q->nk;
    in.fft_rho_grid_points = static_cast<long long>(rho[0]) * rho[1] * rho[2];
    in.fft_wave_grid_points = static_cast<long long>(wave[0]) * wave[1] * wave[2];
    in.davidson_basis_expansion_factor = q->options.basis_factor >= 1 ? q->options.basis_factor : 2;
    in.overlap_is_identity = 1;
    q3x_memory_estimate_compute(&in, &f.ref);
    static const char* const held[] = {"Davidson basis V, HV", "rotation scratch", "subspace matrices",
                                       "nonlocal projectors", "block-diagonal D matrix"};
    for (int i = 0; i < f.ref.n_steady_state; ++i) {
        const q3x_memory_term& t = f.ref.steady_state[i];
        if (std::strcmp(t.label, "wavefunctions") == 0) f.psi += static_cast<double>(t.bytes);
        for (const char* h : held)
            if (std::strcmp(t.label, h) == 0) f.per_kpoint += static_cast<double>(t.bytes);
    }
    const double m = f.ref.maximum_basis_size, n = npw_est, nb = q->nbands, n3 = wave[2];
    const double sticks = std::max(1.0, n / std::max(1.0, n3)) * 1.6;  // occupied z-sticks, with margin
    auto dev = [&](const char* label, const char* formula, double bytes) {
        const unsigned long long b = static_cast<unsigned long long>(std::ceil(bytes));
        f.dev[f.ndev++] = {label, formula, b};
        f.per_kpoint += static_cast<double>(b);
    };
    dev("second Davidson basis V, HV", "2 x M x npw x 16", 2.0 * m * n * 16.0);
    dev("subspace eigensolver and Gram workspace", "26 x M^2 x 16", 26.0 * m * m * 16.0);
    dev("nonlocal projectors, complex64 copy", "nkb x npw x 8", static_cast<double>(nproj) * n * 8.0);
    dev("z-stick buffer", "nb x N3 x 1.6 max(1, npw / N3) x 16", nb * n3 * sticks * 16.0);
    dev("potential, fp64 and fp32", "Nwave x 12", 12.0 * static_cast<double>(in.fft_wave_grid_points));
    f.total = f.psi + f.per_kpoint * std::min(q->nk, concurrent);
    return Q3X_OK;
}

}  // namespace

extern "C" {

q3x_status q3x_server_create(int device, int workers, double hbm_budget_bytes, q3x_server* out) {
    try {
        if (!out || workers < 1 || workers > 64) {
            set_error("q3x_server_create: need 1 <= workers <= 64");
            return Q3X_ERR_ARGUMENT;
        }
        auto* s = new q3x_server_s;
        s->device = device;
        s->t0 = clk::now();
        for (int w = 0; w < workers; ++w) {
            q3x_context c = nullptr;
            const q3x_status st = q3x_context_create(device, &c);
            if (st != Q3X_OK) {
                for (q3x_context x : s->workers) q3x_context_destroy(x);
                delete s;
                return st;
            }
            s->workers.push_back(c);
        }
        size_t free_b = 0, total_b = 0;
        (void)hipMemGetInfo(&free_b, &total_b);
        s->budget = hbm_budget_bytes > 0 ? hbm_budget_bytes : 0.8 * static_cast<double>(free_b);
        s->hams.assign(workers, nullptr);
        s->worker_bytes.assign(workers, 0.0);
        s->ham_ticket.assign(workers, 0);
        // host preparation runs ahead on half as many threads as workers (its share of a k-point's
        // time is about a third of the device part at the cfg5 cell sizes)
        const int nprep = std::max(2, (workers + 1) / 2);
        s->max_ready = static_cast<size_t>(2 * workers);
        for (int w = 0; w < workers; ++w) s->threads.emplace_back(&q3x_server_s::worker, s, w);
        for (int i = 0; i < nprep; ++i) s->prep_threads.emplace_back(&q3x_server_s::preparer, s);
        *out = s;
        return Q3X_OK;
    } catch (const std::exception& e) {  // nothing crosses the C ABI as an exception
        set_error(std::string("q3x_server_create: ") + e.what());
        return Q3X_ERR_INTERNAL;
    }
}

q3x_status q3x_server_destroy(q3x_server s) {
    try {
        if (!s) return Q3X_OK;
        {
            std::unique_lock<std::mutex> g(s->mu);
            s->cv_done.wait(g, [&] {
                for (auto& kv : s->all)
                    if (!kv.second->done) return false;
                return true;
            });
            s->stop = true;
            s->cv_task.notify_all();
            s->cv_ready.notify_all();
        }
        for (auto& t : s->prep_threads) t.join();
        for (auto& t : s->threads) t.join();
        for (auto& kv : s->all)
            for (auto& cp : kv.second->psi)
                if (cp.second) q3x_device_free(cp.first, cp.second);
        for (auto& kv : s->psi_pool) q3x_device_free(s->workers[0], kv.second);
        for (q3x_hamiltonian h : s->hams)
            if (h) q3x_hamiltonian_destroy(h);
        for (q3x_context c : s->workers) q3x_context_destroy(c);
        delete s;
        return Q3X_OK;
    } catch (const std::exception& e) {  // nothing crosses the C ABI as an exception
        set_error(std::string("q3x_server_destroy: ") + e.what());
        return Q3X_ERR_INTERNAL;
    }
}

q3x_status q3x_server_cancel_pending(q3x_server s, int* cancelled) {
    try {
        if (!s) {
            set_error("q3x_server_cancel_pending: null server");
            return Q3X_ERR_ARGUMENT;
        }
        std::lock_guard<std::mutex> g(s->mu);
        int n = 0;
        auto finish = [&](Request& r) {  // k-points still running finish on their own
            if (r.status == Q3X_OK) {
                r.status = Q3X_ERR_STATE;
                r.error = "cancelled before all k-points were solved";
                ++n;
            }
            if (r.remaining == 0 && !r.done) {
                r.done = true;
                r.t_done = now_s(s->t0);
                if (r.admitted) {
                    s->work_running -= r.bytes - r.bytes_psi;
                    if (s->work_running < 1.0) s->work_running = 0.0;
                }
            }
        };
        // not admitted: nothing of theirs is on the device or in the budget
        for (auto& r : s->waiting) {
            r->remaining = 0;
            finish(*r);
        }
        s->waiting.clear();
        // admitted, k-points not yet handed to a preparation thread
        for (auto& r : s->running) {
            r->remaining -= r->kpath ? (r->next_task < r->nk ? 1 : 0) : r->nk - r->next_task;
            r->next_task = r->nk;
            finish(*r);
        }
        s->running.clear();
        // prepared, not yet taken by a worker
        for (auto& p : s->ready) {
            if (p->kp) q3x_kpoint_destroy(p->kp);
            --p->r->remaining;
            finish(*p->r);
        }
        s->ready.clear();
        // k-path requests on a worker (or on their way to one) stop before their next solve
        for (auto& kv : s->all) {
            Request& r = *kv.second;
            if (r.kpath && r.admitted && !r.done) {
                r.cancel = true;
                finish(r);
            }
        }
        if (cancelled) *cancelled = n;
        s->admit_locked();
        s->cv_done.notify_all();
        s->cv_task.notify_all();
        return Q3X_OK;
    } catch (const std::exception& e) {  // nothing crosses the C ABI as an exception
        set_error(std::string("q3x_server_cancel_pending: ") + e.what());
        return Q3X_ERR_INTERNAL;
    }
}

namespace {
q3x_status estimate(q3x_server s, const q3x_band_request* q, double* bytes, double* psi_share) {
    try {
        if (!s || !q || !bytes) {
            set_error("q3x_server_estimate: invalid argument");
            return Q3X_ERR_ARGUMENT;
        }
        // a k-path request runs on one worker at a time
        const int concurrent = q->kpath_propagation ? 1 : static_cast<int>(s->workers.size());
        Footprint f;
        const q3x_status st = request_footprint(q, concurrent, f);
        if (st != Q3X_OK) return st;
        *bytes = f.total;
        if (psi_share) *psi_share = f.psi;
        return Q3X_OK;
    } catch (const std::exception& e) {  // nothing crosses the C ABI as an exception
        set_error(std::string("q3x_server_estimate: ") + e.what());
        return Q3X_ERR_INTERNAL;
    }
}
}  // namespace

q3x_status q3x_server_estimate(q3x_server s, const q3x_band_request* q, double* bytes) {
    return estimate(s, q, bytes, nullptr);
}

q3x_status q3x_band_request_footprint(const q3x_band_request* q, int concurrent_kpoints,
                                      q3x_memory_estimate* reference_out, q3x_memory_term* device_terms,
                                      int* n_device_terms, double* bytes) {
    try {
        Footprint f;
        const q3x_status st = request_footprint(q, concurrent_kpoints, f);
        if (st != Q3X_OK) return st;
        if (reference_out) *reference_out = f.ref;
        if (device_terms)
            for (int i = 0; i < f.ndev; ++i) device_terms[i] = f.dev[i];
        if (n_device_terms) *n_device_terms = f.ndev;
        if (bytes) *bytes = f.total;
        return Q3X_OK;
    } catch (const std::exception& e) {  // nothing crosses the C ABI as an exception
        set_error(std::string("q3x_band_request_footprint: ") + e.what());
        return Q3X_ERR_INTERNAL;
    }
}

q3x_status q3x_server_submit(q3x_server s, const q3x_band_request* q, long* ticket) {
    try {
        if (!s || !q || !q->basis || !q->v_r || !q->kpoints || q->nk < 1 || q->nbands < 1 || !ticket ||
            q->kpath_propagation < 0 || q->kpath_propagation > 1 || q->band_path_segments < 0 ||
            q->band_path_points_per_segment < 0) {
            set_error("q3x_server_submit: invalid argument");
            return Q3X_ERR_ARGUMENT;
        }
        auto r = std::make_shared<Request>();
        const q3x_status st = estimate(s, q, &r->bytes, &r->bytes_psi);
        if (st != Q3X_OK) return st;
        r->basis = q->basis;
        r->v = copy_of(q->v_r, static_cast<size_t>(q->nv));
        r->kpts = copy_of(q->kpoints, static_cast<size_t>(q->nk) * 3);
        r->nk = q->nk;
        r->natoms = q->natoms;
        r->nspecies = q->nspecies;
        r->atom_pos = copy_of(q->atom_pos, static_cast<size_t>(q->natoms) * 3);
        r->atom_species = copy_of(q->atom_species, static_cast<size_t>(q->natoms));
        r->nmesh = copy_of(q->nmesh, static_cast<size_t>(q->nspecies));
        r->nbeta = copy_of(q->nbeta, static_cast<size_t>(q->nspecies));
        const size_t mesh_total = std::accumulate(r->nmesh.begin(), r->nmesh.end(), size_t(0));
        size_t beta_total = 0, beta_values = 0;
        for (int sp = 0; sp < q->nspecies; ++sp) {
            beta_total += r->nbeta[sp];
            beta_values += static_cast<size_t>(r->nbeta[sp]) * r->nmesh[sp];
        }
        r->r_all = copy_of(q->r_all, mesh_total);
        r->rab_all = copy_of(q->rab_all, mesh_total);
        r->l_all = copy_of(q->l_all, beta_total);
        r->cutoff_all = copy_of(q->cutoff_all, beta_total);
        r->beta_r_all = copy_of(q->beta_r_all, beta_values);
        r->block_sizes = copy_of(q->block_sizes, static_cast<size_t>(q->natoms));
        size_t dvals = 0;
        for (int b : r->block_sizes) dvals += static_cast<size_t>(b) * b;
        r->d_blocks = copy_of(q->d_blocks, dvals);
        r->nb = q->nbands;
        r->nocc = q->noccupied;
        r->opt = q->options;
        r->opt.n_occupied = q->noccupied;
        r->kpath = q->kpath_propagation == 1;
        if (r->kpath) r->steps = nscf_schedule(q->nk, q->band_path_segments, q->band_path_points_per_segment);
        r->remaining = r->kpath ? 1 : q->nk;
        r->eig.assign(static_cast<size_t>(q->nk) * q->nbands, 0.0);
        r->psi.assign(q->nk, {nullptr, nullptr});
        r->psi_bytes.assign(q->nk, 0);
        std::lock_guard<std::mutex> g(s->mu);
        r->ticket = s->next_ticket++;
        r->t_submit = now_s(s->t0);
        s->all[r->ticket] = r;
        s->waiting.push_back(r);
        s->admit_locked();
        *ticket = r->ticket;
        return Q3X_OK;
    } catch (const std::exception& e) {  // nothing crosses the C ABI as an exception
        set_error(std::string("q3x_server_submit: ") + e.what());
        return Q3X_ERR_INTERNAL;
    }
}

q3x_status q3x_server_wait(q3x_server s, long ticket, double* eig_out, q3x_request_stats* stats) {
    try {
        if (!s) {
            set_error("q3x_server_wait: null server");
            return Q3X_ERR_ARGUMENT;
        }
        std::shared_ptr<Request> r;
        {
            std::unique_lock<std::mutex> g(s->mu);
            auto it = s->all.find(ticket);
            if (it == s->all.end()) {
                set_error("q3x_server_wait: unknown ticket");
                return Q3X_ERR_ARGUMENT;
            }
            r = it->second;
            s->cv_done.wait(g, [&] { return r->done; });
            // erase by key: another wait on the same ticket may have collected it meanwhile
            auto again = s->all.find(ticket);
            if (again == s->all.end() || again->second != r) {
                set_error("q3x_server_wait: the ticket was collected by another call");
                return Q3X_ERR_ARGUMENT;
            }
            s->all.erase(again);
        }
        if (eig_out) std::memcpy(eig_out, r->eig.data(), r->eig.size() * sizeof(double));
        if (stats) {
            stats->submit_s = r->t_submit;
            stats->admit_s = r->t_admit;
            stats->done_s = r->t_done;
            stats->footprint_bytes = r->bytes;
            stats->iterations = r->iterations;
            stats->setup_s = r->setup_s;
            stats->solve_s = r->solve_s;
            stats->eig_barrier_redos = r->eig_barrier_redos;
            stats->kpoints_done = r->kpoints_done;
            stats->solves = r->solves;
        }
        {
            // the request's wavefunction blocks go back to the pool, and its share of the
            // admission budget moves with them (the pool is freed when admission needs the room)
            std::lock_guard<std::mutex> g(s->mu);
            for (size_t k = 0; k < r->psi.size(); ++k)
                if (r->psi[k].second) {
                    s->psi_pool.emplace(r->psi_bytes[k], r->psi[k].second);
                    s->pool_bytes += static_cast<double>(r->psi_bytes[k]);
                }
            if (r->admitted) s->psi_held -= r->bytes_psi;  // a request cancelled before admission held none
            if (s->psi_held < 1.0) s->psi_held = 0.0;
            s->admit_locked();
        }
        r->psi.clear();
        if (r->status != Q3X_OK) set_error(r->error);
        return r->status;
    } catch (const std::exception& e) {  // nothing crosses the C ABI as an exception
        set_error(std::string("q3x_server_wait: ") + e.what());
        return Q3X_ERR_INTERNAL;
    }
}

double q3x_server_in_flight_bytes(q3x_server s) {
    if (!s) return 0.0;
    std::lock_guard<std::mutex> g(s->mu);
    return s->resident_locked();
}

}  // extern "C"
