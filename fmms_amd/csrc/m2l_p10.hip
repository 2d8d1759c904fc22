// m2l_p10.hip — M2L contraction kernels for p = 10 (see m2l_kernel.h).
#include "m2l_kernel.h"

namespace fmmh {
FMMH_M2L_INSTANTIATE(10)
}  // namespace fmmh
