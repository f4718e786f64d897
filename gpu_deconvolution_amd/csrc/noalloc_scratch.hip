// Device scratch of the allocation-free RL entry point
// (fmvd_deconvolve_volume_device_noalloc_cuda_*): a file-scope device array
// that every GPU the library runs on holds from module load until the
// process ends.  Host code never passes the array's name to a kernel (in
// host code it names a host-side stand-in): the first call on each device
// takes the device address with hipGetSymbolAddress, with that device
// current, and every later use goes through that pointer.
#include <hip/hip_runtime.h>

#include <mutex>

#include "hip_check.hpp"
#include "noalloc_scratch.hpp"

namespace fmvd {

static __device__ __align__(256) unsigned char scratch[kNoAllocScratchBytes];

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_mu;
void* g_ptr[kMaxDevices] = {};
}  // namespace

void* noalloc_scratch()
{
    int dev = 0;
    FMVD_HIP(hipGetDevice(&dev));
    if (dev < 0 || dev >= kMaxDevices) throw_msg("device index out of range for the scratch table");
    std::lock_guard<std::mutex> lk(g_mu);
    if (!g_ptr[dev]) {
        void* p = nullptr;
        FMVD_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        if (!p) throw_msg("hipGetSymbolAddress returned no address for the scratch array");
        g_ptr[dev] = p;
    }
    return g_ptr[dev];
}

}  // namespace fmvd
