// Instantiation unit: uneven-batch persistent kernel, large class: 16 values per thread, 1024 threads, complex128.
// (One kernel per unit: the units compile in parallel.)
#include "kernels_body.h"

namespace fftz {

const void* kernel_symbol_batch_f64_large() { return (const void*)&k_batch<double, 16, 1024>; }

}  // namespace fftz
