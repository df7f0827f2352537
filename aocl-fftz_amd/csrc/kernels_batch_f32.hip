// Instantiation unit: uneven-batch persistent kernels (small / medium / large classes), complex64.
#include "kernels_body.h"

namespace fftz {

const void* kernel_symbol_batch_f32(int E, int maxt)
{
    if (E == 8) return (const void*)&k_batch<float, 8, 256>;
    return maxt == 512 ? (const void*)&k_batch<float, 16, 512> : (const void*)&k_batch<float, 16, 1024>;
}

}  // namespace fftz
