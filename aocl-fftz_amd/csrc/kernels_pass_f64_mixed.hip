// Instantiation unit: single-size pass and Bluestein kernels with every radix (2-13),
// complex128; 8 or 16 values per thread, 256- or 1024-thread bound.
#include "kernels_body.h"

namespace fftz {

const void* kernel_symbol_pass_f64_mixed(KernelKind kind, int E, int maxt)
{
    return sym_et<double, true>(kind, E, maxt);
}

}  // namespace fftz
