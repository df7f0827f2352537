// Launch table of the LDS pass kernels: precision x radix set x elements per
// thread x block limit.  The kernels themselves are instantiated in
// kernels_pass_{f32,f64}_{mixed,pow2}.hip and
// kernels_batch_{f32,f64}_{small,medium,large}.hip (one translation unit
// each, so they compile in parallel) from kernels_body.h.
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
    if (prec_bytes == 4)
        return all_radix ? kernel_symbol_pass_f32_mixed(kind, E, maxt) : kernel_symbol_pass_f32_pow2(kind, E, maxt);
    return all_radix ? kernel_symbol_pass_f64_mixed(kind, E, maxt) : kernel_symbol_pass_f64_pow2(kind, E, maxt);
}

const void* kernel_symbol_batch(int prec_bytes, int E, int maxt)
{
    if (E == 8) return prec_bytes == 4 ? kernel_symbol_batch_f32_small() : kernel_symbol_batch_f64_small();
    if (maxt == 512) return prec_bytes == 4 ? kernel_symbol_batch_f32_medium() : kernel_symbol_batch_f64_medium();
    return prec_bytes == 4 ? kernel_symbol_batch_f32_large() : kernel_symbol_batch_f64_large();
}

const void* kernel_symbol_reset() { return (const void*)&k_reset_counters; }

}  // namespace fftz
