// Class plans (21|10), (21|11), (21|20), (22|00), (22|10), (22|20): instantiations of the register kernels of
// jk_fast_impl.h, compiled as a separate translation unit.
// (la, lb, lc, ld, chunks over A-components, chunks over C-components); the
// splits keep contracted intermediates + tile block + VRR tables within about
// 110 doubles per lane (no scratch, >= 2 waves/SIMD on gfx950).
#include "jk_fast_impl.h"

#define ASUKA_PLANS_P1(X) \
  X(2, 1, 1, 0, 1, 2) \
  X(2, 1, 1, 1, 6, 1) \
  X(2, 1, 2, 0, 1, 3) \
  X(2, 2, 0, 0, 1, 1) \
  X(2, 2, 1, 0, 1, 3) \
  X(2, 2, 2, 0, 3, 3) \

ASUKA_FAST_FIND(find_launcher_p1, ASUKA_PLANS_P1)
