// Dispatch to the lane-per-quartet derivative kernels (jk_grad_fast_impl.h);
// the class instantiations are spread over jk_grad_fast_p*.hip so they compile
// in parallel.
#include "jk_grad_fast_impl.h"

asuka_gradf::DerivLauncher find_gradf_p0(int la, int lb, int lc, int ld);
asuka_gradf::DerivLauncher find_gradf_p1(int la, int lb, int lc, int ld);
asuka_gradf::DerivLauncher find_gradf_p2(int la, int lb, int lc, int ld);
asuka_gradf::DerivLauncher find_gradf_p3(int la, int lb, int lc, int ld);

hipError_t launch_eri_deriv_fast(const DevBasis& b, const ClassWork& w, const double* D, const double* WJ,
                                 const double* WK, const int* shell_atom, int natm, double* grad, hipStream_t s,
                                 bool* handled) {
  *handled = false;
  if (!w.units || !w.tasks) return hipSuccess;
  asuka_gradf::DerivLauncher f = nullptr;
  asuka_gradf::DerivLauncher (*finders[])(int, int, int, int) = {find_gradf_p0, find_gradf_p1, find_gradf_p2,
                                                                  find_gradf_p3};
  for (auto fn : finders)
    if (!f) f = fn(w.la, w.lb, w.lc, w.ld);
  if (!f) return hipSuccess;
  *handled = true;
  if (w.ntasks <= 0) return hipSuccess;
  asuka_gradf::GradFastArgs g{D, WJ, WK, shell_atom, grad, natm};
  return f(b, w, g, s);
}
