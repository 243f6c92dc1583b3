// Derivative-kernel classes (0, 0, 0, 0, 1), (1, 0, 0, 0, 1), (1, 0, 1, 0, 1), (1, 1, 0, 0, 1), (2, 0, 0, 0, 1), (3, 0, 0, 0, 1), (1, 1, 1, 0, 3) (la, lb, lc, ld, component chunks):
// instantiations of jk_grad_fast_impl.h, compiled as a separate translation unit.
#include "jk_grad_fast_impl.h"

#define ASUKA_GRADF_P0(X) \
  X(0, 0, 0, 0, 1) \
  X(1, 0, 0, 0, 1) \
  X(1, 0, 1, 0, 1) \
  X(1, 1, 0, 0, 1) \
  X(2, 0, 0, 0, 1) \
  X(3, 0, 0, 0, 1) \
  X(1, 1, 1, 0, 3) \

ASUKA_GRADF_FIND(find_gradf_p0, ASUKA_GRADF_P0)
