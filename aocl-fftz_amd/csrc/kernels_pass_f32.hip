// Instantiation unit: single-size pass and Bluestein kernels, complex64.
#include "kernels_body.h"

namespace fftz {

const void* kernel_symbol_pass_f32(KernelKind kind, bool all_radix, int E, int maxt)
{
    return all_radix ? sym_et<float, true>(kind, E, maxt) : sym_et<float, false>(kind, E, maxt);
}

}  // namespace fftz
