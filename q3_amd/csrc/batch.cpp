// Many independent k-points (or cells) solved together on one GPU.
//
// The reference solves k-states one after another in the SCF k-loop
// (iterator/kohn_sham/planewave/pseudopotential/init.hpp:1056-1082) and spreads k-points over
// MPI ranks round-robin (iterator/kohn_sham/planewave/internal/kpoint_distributor.hpp).  On one
// MI355X a single small k-point leaves most of the 256 CUs idle (the subspace eigensolver is
// one workgroup, the H-apply of 48 bands a few thousand), so the batch spreads its k-points
// round-robin over worker contexts, each with its own HIP stream, workspace and host thread:
// the workers' kernels run concurrently on the device and each worker's host decisions
// (convergence tests, corrections) overlap the others' device work.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <mutex>
#include <stdexcept>
#include <string>
#include <thread>
#include <vector>

#include "runtime.h"

struct q3x_batch_s {
    int device = 0;
    std::vector<q3x_context> workers;
    std::vector<q3x_hamiltonian> hams;  // k-point i lives on workers[i % workers.size()]
    std::vector<int> npw;
};

using namespace q3x;

extern "C" {

q3x_status q3x_batch_create(int device, int workers, q3x_batch* out) {
    try {
        if (!out || workers < 1 || workers > 64) {
            set_error("q3x_batch_create: need 1 <= workers <= 64");
            return Q3X_ERR_ARGUMENT;
        }
        auto* b = new q3x_batch_s;
        b->device = device;
        for (int w = 0; w < workers; ++w) {
            q3x_context c = nullptr;
            const q3x_status s = q3x_context_create(device, &c);
            if (s != Q3X_OK) {
                q3x_batch_destroy(b);
                return s;
            }
            b->workers.push_back(c);
        }
        *out = b;
        return Q3X_OK;
    } catch (const std::exception& e) {  // nothing crosses the C ABI as an exception
        set_error(std::string("q3x_batch_create: ") + e.what());
        return Q3X_ERR_INTERNAL;
    }
}

q3x_status q3x_batch_destroy(q3x_batch b) {
    try {
        if (!b) return Q3X_OK;
        for (q3x_hamiltonian h : b->hams) q3x_hamiltonian_destroy(h);
        for (q3x_context c : b->workers) q3x_context_destroy(c);
        delete b;
        return Q3X_OK;
    } catch (const std::exception& e) {  // nothing crosses the C ABI as an exception
        set_error(std::string("q3x_batch_destroy: ") + e.what());
        return Q3X_ERR_INTERNAL;
    }
}

q3x_status q3x_batch_add(q3x_batch b, q3x_basis basis, q3x_kpoint kp, const double* v_r, long n,
                         const double* beta, int nproj, int natoms, const int* block_sizes,
                         const double* d_blocks, int* index) {
    try {
        if (!b || !basis || !kp || !v_r) {
            set_error("q3x_batch_add: invalid argument");
            return Q3X_ERR_ARGUMENT;
        }
        const int i = static_cast<int>(b->hams.size());
        q3x_context ctx = b->workers[i % b->workers.size()];
        q3x_hamiltonian h = nullptr;
        q3x_status s = q3x_hamiltonian_create(ctx, basis, &h);
        if (s != Q3X_OK) return s;
        if ((s = q3x_hamiltonian_set_potential(h, v_r, n)) != Q3X_OK ||
            (s = q3x_hamiltonian_prepare_kpoint(h, kp)) != Q3X_OK ||
            (nproj > 0 &&
             (s = q3x_hamiltonian_set_projectors(h, beta, nproj, natoms, block_sizes, d_blocks)) != Q3X_OK)) {
            q3x_hamiltonian_destroy(h);
            return s;
        }
        b->hams.push_back(h);
        b->npw.push_back(h->npw);
        if (index) *index = i;
        return Q3X_OK;
    } catch (const std::exception& e) {  // nothing crosses the C ABI as an exception
        set_error(std::string("q3x_batch_add: ") + e.what());
        return Q3X_ERR_INTERNAL;
    }
}

int q3x_batch_size(q3x_batch b) { return b ? static_cast<int>(b->hams.size()) : 0; }

q3x_status q3x_batch_solve(q3x_batch b, int nb, const q3x_davidson_options* opt, double* const* psi_inout,
                           double* eig_out, q3x_davidson_report* reports) {
    try {
        if (!b || !opt || !psi_inout || !eig_out || nb <= 0) {
            set_error("q3x_batch_solve: invalid argument");
            return Q3X_ERR_ARGUMENT;
        }
        const int nk = static_cast<int>(b->hams.size());
        const int nw = static_cast<int>(b->workers.size());
        std::mutex mu;
        q3x_status first = Q3X_OK;
        std::string first_msg;
        std::atomic<bool> failed{false};
        auto work = [&](int w) {
            if (hipSetDevice(b->device) != hipSuccess) {
                std::lock_guard<std::mutex> g(mu);
                if (first == Q3X_OK) {
                    first = Q3X_ERR_DEVICE;
                    first_msg = "q3x_batch_solve: hipSetDevice failed in a worker";
                }
                failed = true;
                return;
            }
            for (int i = w; i < nk && !failed; i += nw) {
                q3x_davidson_report rep{};
                const q3x_status s = q3x_davidson_solve(b->hams[i], nb, opt, psi_inout[i],
                                                        eig_out + static_cast<size_t>(i) * nb, &rep);
                if (reports) reports[i] = rep;
                if (s != Q3X_OK) {
                    std::lock_guard<std::mutex> g(mu);
                    if (first == Q3X_OK) {
                        first = s;
                        first_msg = "q3x_batch_solve: k-point " + std::to_string(i) + ": " + q3x_last_error();
                    }
                    failed = true;
                }
            }
        };
        std::vector<std::thread> threads;
        for (int w = 1; w < nw; ++w) threads.emplace_back(work, w);
        work(0);
        for (auto& t : threads) t.join();
        if (first != Q3X_OK) set_error(first_msg);
        return first;
    } catch (const std::exception& e) {  // nothing crosses the C ABI as an exception
        set_error(std::string("q3x_batch_solve: ") + e.what());
        return Q3X_ERR_INTERNAL;
    }
}

}  // extern "C"
