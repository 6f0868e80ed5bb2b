// mttkrp_f32.hip -- MTTKRP kernels with float factors (float values), instantiated
// for 2..5 modes and every column layout.
#include "mttkrp_kern.hpp"

namespace blco {

const void* mttkrp_kernel_f32(int N, const Shape& s, int which)
{
    return select_n<float, float>(N, s, which);
}

}  // namespace blco
