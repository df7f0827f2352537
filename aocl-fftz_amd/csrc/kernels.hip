// Launch table of the LDS pass kernels: precision x radix set x elements per
// thread x block limit.  The kernels themselves are instantiated in
// kernels_pass_{f32,f64}.hip and kernels_batch_{f32,f64}.hip (one translation
// unit each, so they compile in parallel) from kernels_body.h.
#include <hip/hip_runtime.h>

#include "kernels.h"

namespace fftz {

// Clears the uneven batch's work-queue counters (a kernel rather than a
// memset node, so the launch sequence can be captured in a graph).
__global__ void k_reset_counters(unsigned* c, int n)
{
    if ((int)threadIdx.x < n) c[threadIdx.x] = 0u;
}

const void* kernel_symbol(KernelKind kind, int prec_bytes, bool all_radix, int E, int maxt)
{
    return prec_bytes == 4 ? kernel_symbol_pass_f32(kind, all_radix, E, maxt)
                           : kernel_symbol_pass_f64(kind, all_radix, E, maxt);
}

const void* kernel_symbol_batch(int prec_bytes, int E, int maxt)
{
    return prec_bytes == 4 ? kernel_symbol_batch_f32(E, maxt) : kernel_symbol_batch_f64(E, maxt);
}

const void* kernel_symbol_reset() { return (const void*)&k_reset_counters; }

}  // namespace fftz
