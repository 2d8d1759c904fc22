// m2l_p1_6.hip — M2L contraction kernels for p = 1..6 (see m2l_kernel.h).
#include "m2l_kernel.h"

namespace fmmh {
FMMH_M2L_INSTANTIATE(1)
FMMH_M2L_INSTANTIATE(2)
FMMH_M2L_INSTANTIATE(3)
FMMH_M2L_INSTANTIATE(4)
FMMH_M2L_INSTANTIATE(5)
FMMH_M2L_INSTANTIATE(6)
}  // namespace fmmh
