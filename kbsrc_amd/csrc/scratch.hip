// The library's fixed scratch memory for executes that must not allocate (kbsrc_plan_execute_noalloc): one
// file-scope device array per GPU, held from its first use until the process ends.
//
// In host code the array's name only stands in for the symbol; its device address comes from hipGetSymbolAddress,
// taken once per device with that device current, and only that address is handed to kernels and copies.
#include <hip/hip_runtime.h>

#include <mutex>

#include "engine.hpp"

namespace kbsrc::gpu {

namespace {

__device__ __align__(256) unsigned char scratch[kScratchBytes];

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
void* g_scratch_ptr[kMaxDevices] = {};

}  // namespace

double* scratch_memory() {
    int dev = 0;
    hip_check(hipGetDevice(&dev), "device");
    if (dev < 0 || dev >= kMaxDevices) throw HipError("static scratch: device index out of range");
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
        void* p = nullptr;
        hip_check(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)), "static scratch address");
        if (!p) throw HipError("static scratch address: null");
        g_scratch_ptr[dev] = p;
    }
    return static_cast<double*>(g_scratch_ptr[dev]);
}

}  // namespace kbsrc::gpu
