// Device scratch of the allocation-free execute (aoclfftz_gpu_execute_noalloc).
//
// The array below is a file-scope __device__ variable: the code-object loader
// places one copy of it on every GPU that loads this library, and it stays
// there until the process ends.  In host code the name `scratch` is only a
// host-side stand-in for that variable, not a device address; the device
// address of each GPU's copy comes from hipGetSymbolAddress, called once per
// device with that device current, and is the only form host code passes to
// kernels.
#include <hip/hip_runtime.h>

#include <mutex>

#include "noalloc.h"

namespace fftz {

// [0, kNoAllocHeaderBytes): uneven-batch work-queue counters (cleared by a
// kernel at the start of every execute); then the work area (four-step and
// multi-pass intermediates, fully written before they are read).
static __device__ __align__(256) unsigned char scratch[kNoAllocHeaderBytes + kNoAllocWorkBytes];

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_mu;
void* g_ptr[kMaxDevices] = {};
}  // namespace

hipError_t noalloc_scratch(int device, void** ptr)
{
    *ptr = nullptr;
    if (device < 0 || device >= kMaxDevices) return hipErrorInvalidDevice;
    std::lock_guard<std::mutex> g(g_mu);
    if (!g_ptr[device]) {
        void* p = nullptr;
        hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess) return e;
        g_ptr[device] = p;
    }
    *ptr = g_ptr[device];
    return hipSuccess;
}

}  // namespace fftz
