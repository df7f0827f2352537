// Instantiation unit: single-size pass and Bluestein kernels, complex128.
#include "kernels_body.h"

namespace fftz {

const void* kernel_symbol_pass_f64(KernelKind kind, bool all_radix, int E, int maxt)
{
    return all_radix ? sym_et<double, true>(kind, E, maxt) : sym_et<double, false>(kind, E, maxt);
}

}  // namespace fftz
