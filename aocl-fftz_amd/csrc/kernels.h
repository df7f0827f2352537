// Kernel symbols exported from the kernel units to the host planner.
#pragma once

#include "pass.h"

namespace fftz {

enum class KernelKind { Pass, Bluestein };

// Device function symbol for hipLaunchKernel / hipFuncSetAttribute.
// prec_bytes: 4 (complex64) or 8 (complex128).  all_radix: the variant that
// carries the 3/5/6/7/9/10/11/12/13 stages (the power-of-two variant carries
// only 2/4/8/16 and therefore needs fewer VGPRs).  E: complex values per
// thread per stage (8 or 16).  maxt: block-size bound of the variant (256 or
// 1024); the 256 variants get a larger VGPR budget.
const void* kernel_symbol(KernelKind kind, int prec_bytes, bool all_radix, int E, int maxt);

// Uneven-batch persistent kernel (k_batch) and the counter reset kernel
// launched before it.  Classes: E = 8 with 256 threads, E = 16 with 512 or
// 1024 threads.
const void* kernel_symbol_batch(int prec_bytes, int E, int maxt);
const void* kernel_symbol_reset();

// Instantiation units behind kernel_symbol / kernel_symbol_batch (one
// translation unit each, compiled in parallel).
const void* kernel_symbol_pass_f32_mixed(KernelKind kind, int E, int maxt);
const void* kernel_symbol_pass_f32_pow2(KernelKind kind, int E, int maxt);
const void* kernel_symbol_pass_f64_mixed(KernelKind kind, int E, int maxt);
const void* kernel_symbol_pass_f64_pow2(KernelKind kind, int E, int maxt);
const void* kernel_symbol_batch_f32_small();
const void* kernel_symbol_batch_f32_medium();
const void* kernel_symbol_batch_f32_large();
const void* kernel_symbol_batch_f64_small();
const void* kernel_symbol_batch_f64_medium();
const void* kernel_symbol_batch_f64_large();

// Fixed-length pass kernel for power-of-two L in [64, 4096] (radix plan 16s
// then remainder, 16 elements per thread), block <= maxt (256, or 512 for
// L >= 256); nullptr otherwise.
const void* kernel_symbol_fixed(int prec_bytes, int L, int maxt);

// Matrix-core pass (kernels_mfma.hip) for complex64, L = 256 or 4096, with
// operand precision `mode` (MfmaMode) and nb (2 or 4) column blocks per
// wavefront (block = 1024 / nb threads); nullptr for other arguments.
// Kernel arguments: (PassDesc, const float* fragments).
const void* kernel_symbol_mfma(int mode, int L, int nb);

}  // namespace fftz
