// Derivative-kernel classes (2, 0, 1, 0, 2), (2, 1, 0, 0, 2), (2, 2, 0, 0, 4), (3, 0, 1, 0, 3), (2, 0, 2, 0, 4) (la, lb, lc, ld, component chunks):
// instantiations of jk_grad_fast_impl.h, compiled as a separate translation unit.
#include "jk_grad_fast_impl.h"

#define ASUKA_GRADF_P1(X) \
  X(2, 0, 1, 0, 2) \
  X(2, 1, 0, 0, 2) \
  X(2, 2, 0, 0, 4) \
  X(3, 0, 1, 0, 3) \
  X(2, 0, 2, 0, 4) \

ASUKA_GRADF_FIND(find_gradf_p1, ASUKA_GRADF_P1)
