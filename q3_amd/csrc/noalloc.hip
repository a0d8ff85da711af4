// H|psi> without device allocation: every scratch buffer of the product lives in one fixed,
// file-scope device array, for callers that cannot call hipMalloc (code under stream capture,
// serving loops that forbid allocation).
#include <hip/hip_runtime.h>

#include <mutex>
#include <string>

#include "q3x_internal.h"
#include "runtime.h"

namespace q3x {
namespace {

// 64 MiB per GPU, held for the life of the process once the library's code object is loaded
// on that GPU.  Every call on one GPU shares it: overlapping calls on one GPU race.
constexpr size_t kScratchBytes = size_t(64) << 20;
__device__ __align__(256) char scratch[kScratchBytes];

constexpr int kMaxDevices = 64;
std::mutex g_mu;
char* g_ptr[kMaxDevices] = {};  // device address of scratch on each GPU (nullptr: not looked up)

}  // namespace

bool noalloc_scratch_init(int device) {
    if (device < 0 || device >= kMaxDevices) {
        set_error("no-allocation scratch: device index out of range");
        return false;
    }
    std::lock_guard<std::mutex> lk(g_mu);
    if (g_ptr[device]) return true;
    int cur = -1;
    if (!hip_ok(hipGetDevice(&cur), "hipGetDevice")) return false;
    if (cur != device && !hip_ok(hipSetDevice(device), "hipSetDevice")) return false;
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (cur != device) (void)hipSetDevice(cur);
    if (!hip_ok(e, "hipGetSymbolAddress(scratch)")) return false;
    g_ptr[device] = static_cast<char*>(p);
    return true;
}

}  // namespace q3x

using namespace q3x;

extern "C" {

unsigned long q3x_noalloc_scratch_bytes(void) { return static_cast<unsigned long>(kScratchBytes); }

q3x_status q3x_hamiltonian_apply_noalloc(q3x_hamiltonian h, const void* d_psi, void* d_hpsi, int nvec,
                                         int precision) {
    try {
        if (!h || !h->ctx) {
            set_error("apply_noalloc: null Hamiltonian");
            return Q3X_ERR_ARGUMENT;
        }
        const int dev = h->ctx->device;
        char* p = nullptr;
        {
            std::lock_guard<std::mutex> lk(g_mu);
            if (dev >= 0 && dev < kMaxDevices) p = g_ptr[dev];
        }
        if (!p) {  // the context's creation looks it up; a failure there is reported here
            if (!noalloc_scratch_init(dev)) return Q3X_ERR_DEVICE;
            std::lock_guard<std::mutex> lk(g_mu);
            p = g_ptr[dev];
        }
        return hamiltonian_apply_fixed(h, static_cast<const double*>(d_psi), static_cast<double*>(d_hpsi), nvec,
                                       precision, p, kScratchBytes);
    } catch (const std::exception& e) {  // nothing crosses the C ABI as an exception
        set_error(std::string("q3x_hamiltonian_apply_noalloc: ") + e.what());
        return Q3X_ERR_INTERNAL;
    }
}

}  // extern "C"
