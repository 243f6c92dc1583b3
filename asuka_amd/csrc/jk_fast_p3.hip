// Class plans (30|00), (30|10), (30|11), (30|20), (30|21), (30|30), (31|00), (31|10), (31|11): instantiations of the register kernels of
// jk_fast_impl.h, compiled as a separate translation unit.
// (la, lb, lc, ld, chunks over A-components, chunks over C-components); the
// splits keep contracted intermediates + tile block + VRR tables within about
// 110 doubles per lane (no scratch, >= 2 waves/SIMD on gfx950).
#include "jk_fast_impl.h"

#define ASUKA_PLANS_P3(X) \
  X(3, 0, 0, 0, 1, 1) \
  X(3, 0, 1, 0, 1, 1) \
  X(3, 0, 1, 1, 3, 1) \
  X(3, 0, 2, 0, 1, 2) \
  X(3, 0, 2, 1, 4, 2) \
  X(3, 0, 3, 0, 1, 4) \
  X(3, 1, 0, 0, 1, 1) \
  X(3, 1, 1, 0, 1, 3) \
  X(3, 1, 1, 1, 10, 1) \

ASUKA_FAST_FIND(find_launcher_p3, ASUKA_PLANS_P3)
