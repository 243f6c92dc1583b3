// Class plans (21|21), (22|11): instantiations of the register kernels of
// jk_fast_impl.h, compiled as a separate translation unit.
// (la, lb, lc, ld, chunks over A-components, chunks over C-components); the
// splits keep contracted intermediates + tile block + VRR tables within about
// 110 doubles per lane (no scratch, >= 2 waves/SIMD on gfx950).
#include "jk_fast_impl.h"

#define ASUKA_PLANS_P2(X) \
  X(2, 1, 2, 1, 3, 6) \
  X(2, 2, 1, 1, 6, 3) \

ASUKA_FAST_FIND(find_launcher_p2, ASUKA_PLANS_P2)
