// Derivative-kernel classes (1, 1, 1, 1, 9), (2, 0, 1, 1, 6), (2, 1, 1, 0, 6), (2, 2, 1, 0, 12) (la, lb, lc, ld, component chunks):
// instantiations of jk_grad_fast_impl.h, compiled as a separate translation unit.
#include "jk_grad_fast_impl.h"

#define ASUKA_GRADF_P2(X) \
  X(1, 1, 1, 1, 9) \
  X(2, 0, 1, 1, 6) \
  X(2, 1, 1, 0, 6) \
  X(2, 2, 1, 0, 12) \

ASUKA_GRADF_FIND(find_gradf_p2, ASUKA_GRADF_P2)
