// Class plans (31|30), (32|20), (33|10): instantiations of the register kernels of
// jk_fast_impl.h, compiled as a separate translation unit.
// (la, lb, lc, ld, chunks over A-components, chunks over C-components); the
// splits keep contracted intermediates + tile block + VRR tables within about
// 110 doubles per lane (no scratch, >= 2 waves/SIMD on gfx950).
#include "jk_fast_impl.h"

#define ASUKA_PLANS_P6(X) \
  X(3, 1, 3, 0, 3, 5) \
  X(3, 2, 2, 0, 3, 6) \
  X(3, 3, 1, 0, 5, 3) \

ASUKA_FAST_FIND(find_launcher_p6, ASUKA_PLANS_P6)
