// Class plans (31|20), (32|00), (32|10), (33|00), (30|22): instantiations of the register kernels of
// jk_fast_impl.h, compiled as a separate translation unit.
// (la, lb, lc, ld, chunks over A-components, chunks over C-components); the
// splits keep contracted intermediates + tile block + VRR tables within about
// 110 doubles per lane (no scratch, >= 2 waves/SIMD on gfx950).
#include "jk_fast_impl.h"

#define ASUKA_PLANS_P4(X) \
  X(3, 1, 2, 0, 1, 6) \
  X(3, 2, 0, 0, 2, 1) \
  X(3, 2, 1, 0, 2, 3) \
  X(3, 3, 0, 0, 4, 1) \
  X(3, 0, 2, 2, 10, 2) \

ASUKA_FAST_FIND(find_launcher_p4, ASUKA_PLANS_P4)
