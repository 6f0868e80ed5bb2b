// mttkrp_f64.hip -- MTTKRP kernels with double factors (double values), instantiated
// for 2..5 modes and every column layout.
#include "mttkrp_kern.hpp"

namespace blco {

const void* mttkrp_kernel_f64(int N, const Shape& s, int which)
{
    return select_n<double, double>(N, s, which);
}

}  // namespace blco
