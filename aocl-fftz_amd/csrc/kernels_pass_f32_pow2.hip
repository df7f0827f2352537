// Instantiation unit: single-size pass and Bluestein kernels with radices 2, 4, 8, 16 only,
// complex64; 8 or 16 values per thread, 256- or 1024-thread bound.
#include "kernels_body.h"

namespace fftz {

const void* kernel_symbol_pass_f32_pow2(KernelKind kind, int E, int maxt)
{
    return sym_et<float, false>(kind, E, maxt);
}

}  // namespace fftz
