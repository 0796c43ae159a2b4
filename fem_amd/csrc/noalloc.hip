// lina_dev_dot without allocation: the same kernels as lina_dev_dot (block
// partials by k_dot_part, their deterministic sum by k_finalize), with the
// partials and the result in a file-scope device array instead of buffers the
// library allocates.  For callers that cannot allocate device memory.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>
#include <mutex>
#include <stdexcept>

#include "../../include/lina_hip.h"
#include "kernels.hpp"
#include "lina_internal.hpp"
#include "solver.hpp"

namespace {

// k_dot_part launches at most DOT_PART_MAX (1024) blocks, one partial each;
// the result follows the partials.  8,256 bytes per device, held until the
// process ends.
constexpr int kPartials = lina::dev::DOT_PART_MAX;
constexpr int kScratchDoubles = kPartials + 8;
static __device__ __align__(256) double scratch[kScratchDoubles];

constexpr int kMaxDevices = 64;
double* g_ptr[kMaxDevices];  // scratch's device address on each device
std::mutex g_mu;

// scratch's address on the current device, looked up once per device
double* scratch_ptr() {
    int dev = -1;
    lina::hip_check(hipGetDevice(&dev), "hipGetDevice");
    if (dev < 0 || dev >= kMaxDevices) throw std::runtime_error("device index beyond the scratch table");
    std::lock_guard<std::mutex> lk(g_mu);
    if (!g_ptr[dev]) {
        void* p = nullptr;
        lina::hip_check(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)), "hipGetSymbolAddress(scratch)");
        g_ptr[dev] = static_cast<double*>(p);
    }
    return g_ptr[dev];
}

}  // namespace

extern "C" double lina_dev_dot_noalloc(const double* a, const double* b, long n) {
    try {
        if (n < 0 || n > INT32_MAX || (n > 0 && (!a || !b)))
            throw std::invalid_argument("lina_dev_dot_noalloc: n must lie in [0, 2^31-1] with non-null vectors");
        double* ptr = scratch_ptr();
        const int np = lina::dev::dot_part<double>(nullptr, (int)n, a, b, ptr);  // np <= kPartials
        lina::dev::finalize(nullptr, np, ptr, ptr + kPartials, 0, 0);
        double out = NAN;
        lina::hip_check(hipMemcpy(&out, ptr + kPartials, sizeof(double), hipMemcpyDeviceToHost), "d2h");
        return out;
    } catch (const std::exception& e) {
        lina::set_error("%s", e.what());
        return NAN;
    }
}
