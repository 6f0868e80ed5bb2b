// mttkrp_bf16.hip -- MTTKRP kernels with bf16 factors (float values), instantiated
// for 2..5 modes and every column layout.
#include "mttkrp_kern.hpp"

namespace blco {

const void* mttkrp_kernel_bf16(int N, const Shape& s, int which)
{
    return select_n<bf16, float>(N, s, which);
}

}  // namespace blco
