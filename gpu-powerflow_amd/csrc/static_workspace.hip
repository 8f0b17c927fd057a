// Fixed device workspace of the allocation-free inner-solver entry point
// (cupf_lu_factor_solve_static): one file-scope device array that every GPU
// the library runs on holds from its first use until the process ends.
//
// In host code the array's name is only a stand-in for the device object;
// its device address is looked up once per device with hipGetSymbolAddress
// (with that device current) and kept here, and that address is what the
// launches receive.
#include "static_workspace.hpp"

#include "hip_common.hpp"

#include <mutex>
#include <stdexcept>
#include <string>

namespace {

__device__ __align__(256) unsigned char scratch[cupf::kStaticWorkspaceBytes];

constexpr int kMaxDevices = 64;
std::mutex g_mu;
void* g_addr[kMaxDevices] = {};

}  // namespace

namespace cupf {

void* static_workspace() {
    int dev = 0;
    CUPF_HIP(hipGetDevice(&dev));
    if (dev < 0 || dev >= kMaxDevices)
        throw std::runtime_error("static workspace: device ordinal " + std::to_string(dev) +
                                 " is beyond the supported " + std::to_string(kMaxDevices));
    std::lock_guard<std::mutex> lock(g_mu);
    if (!g_addr[dev]) {
        void* p = nullptr;
        CUPF_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        if (!p) throw std::runtime_error("static workspace: hipGetSymbolAddress returned null");
        g_addr[dev] = p;
    }
    return g_addr[dev];
}

}  // namespace cupf
