// Instantiation unit: uneven-batch persistent kernel, medium class: 16 values per thread, 512 threads, complex64.
// (One kernel per unit: the units compile in parallel.)
#include "kernels_body.h"

namespace fftz {

const void* kernel_symbol_batch_f32_medium() { return (const void*)&k_batch<float, 16, 512>; }

}  // namespace fftz
