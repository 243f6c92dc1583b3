// Class plans (00|00), (10|00), (10|10), (11|00), (11|10), (11|11), (20|00), (20|10), (20|11), (20|20), (21|00): instantiations of the register kernels of
// jk_fast_impl.h, compiled as a separate translation unit.
// (la, lb, lc, ld, chunks over A-components, chunks over C-components); the
// splits keep contracted intermediates + tile block + VRR tables within about
// 110 doubles per lane (no scratch, >= 2 waves/SIMD on gfx950).
#include "jk_fast_impl.h"

#define ASUKA_PLANS_P0(X) \
  X(0, 0, 0, 0, 1, 1) \
  X(1, 0, 0, 0, 1, 1) \
  X(1, 0, 1, 0, 1, 1) \
  X(1, 1, 0, 0, 1, 1) \
  X(1, 1, 1, 0, 1, 1) \
  X(1, 1, 1, 1, 1, 3) \
  X(2, 0, 0, 0, 1, 1) \
  X(2, 0, 1, 0, 1, 1) \
  X(2, 0, 1, 1, 2, 1) \
  X(2, 0, 2, 0, 1, 1) \
  X(2, 1, 0, 0, 1, 1) \

ASUKA_FAST_FIND(find_launcher_p0, ASUKA_PLANS_P0)
