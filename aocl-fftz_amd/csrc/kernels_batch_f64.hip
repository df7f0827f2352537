// Instantiation unit: uneven-batch persistent kernels (small / medium / large classes), complex128.
#include "kernels_body.h"

namespace fftz {

const void* kernel_symbol_batch_f64(int E, int maxt)
{
    if (E == 8) return (const void*)&k_batch<double, 8, 256>;
    return maxt == 512 ? (const void*)&k_batch<double, 16, 512> : (const void*)&k_batch<double, 16, 1024>;
}

}  // namespace fftz
