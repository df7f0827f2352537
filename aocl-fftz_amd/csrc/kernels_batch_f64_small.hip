// Instantiation unit: uneven-batch persistent kernel, small class: 8 values per thread, 256 threads, complex128.
// (One kernel per unit: the units compile in parallel.)
#include "kernels_body.h"

namespace fftz {

const void* kernel_symbol_batch_f64_small() { return (const void*)&k_batch<double, 8, 256>; }

}  // namespace fftz
