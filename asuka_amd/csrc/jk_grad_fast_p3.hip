// Derivative-kernel classes (2, 1, 1, 1, 18), (2, 1, 2, 0, 12), (2, 1, 2, 1, 36), (2, 2, 1, 1, 36), (2, 2, 2, 0, 24) (la, lb, lc, ld, component chunks):
// instantiations of jk_grad_fast_impl.h, compiled as a separate translation unit.
#include "jk_grad_fast_impl.h"

#define ASUKA_GRADF_P3(X) \
  X(2, 1, 1, 1, 18) \
  X(2, 1, 2, 0, 12) \
  X(2, 1, 2, 1, 36) \
  X(2, 2, 1, 1, 36) \
  X(2, 2, 2, 0, 24) \

ASUKA_GRADF_FIND(find_gradf_p3, ASUKA_GRADF_P3)
