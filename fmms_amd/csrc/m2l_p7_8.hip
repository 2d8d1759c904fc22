// m2l_p7_8.hip — M2L contraction kernels for p = 7..8 (see m2l_kernel.h).
#include "m2l_kernel.h"

namespace fmmh {
FMMH_M2L_INSTANTIATE(7)
FMMH_M2L_INSTANTIATE(8)
}  // namespace fmmh
