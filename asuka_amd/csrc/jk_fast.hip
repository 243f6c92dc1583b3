// Class-specialised register kernels (placeholder until the fast path lands).
#include "jk_internal.h"

hipError_t launch_eri_fast(const DevBasis& b, const ClassWork& w, const EriOut& o, hipStream_t s, bool* handled) {
  (void)b; (void)w; (void)o; (void)s;
  *handled = false;
  return hipSuccess;
}
