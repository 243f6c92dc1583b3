// Class plans (31|21): instantiations of the register kernels of
// jk_fast_impl.h, compiled as a separate translation unit.
// (la, lb, lc, ld, chunks over A-components, chunks over C-components); the
// splits keep contracted intermediates + tile block + VRR tables within about
// 110 doubles per lane (no scratch, >= 2 waves/SIMD on gfx950).
#include "jk_fast_impl.h"

#define ASUKA_PLANS_P5(X) \
  X(3, 1, 2, 1, 5, 6) \

ASUKA_FAST_FIND(find_launcher_p5, ASUKA_PLANS_P5)
