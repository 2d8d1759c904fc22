// m2l_p9.hip — M2L contraction kernels for p = 9 (see m2l_kernel.h).
#include "m2l_kernel.h"

namespace fmmh {
FMMH_M2L_INSTANTIATE(9)
}  // namespace fmmh
