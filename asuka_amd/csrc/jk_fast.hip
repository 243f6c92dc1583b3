// Dispatch to the class-specialised register kernels (jk_fast_impl.h); the
// plans are spread over jk_fast_p*.hip so they compile in parallel.
#include "jk_fast_impl.h"

asuka_fast::Launcher find_launcher_p0(int la, int lb, int lc, int ld);
asuka_fast::Launcher find_launcher_p1(int la, int lb, int lc, int ld);
asuka_fast::Launcher find_launcher_p2(int la, int lb, int lc, int ld);
asuka_fast::Launcher find_launcher_p3(int la, int lb, int lc, int ld);
asuka_fast::Launcher find_launcher_p4(int la, int lb, int lc, int ld);
asuka_fast::Launcher find_launcher_p5(int la, int lb, int lc, int ld);
asuka_fast::Launcher find_launcher_p6(int la, int lb, int lc, int ld);

int find_launcher_p0_nch(int la, int lb, int lc, int ld);
int find_launcher_p1_nch(int la, int lb, int lc, int ld);
int find_launcher_p2_nch(int la, int lb, int lc, int ld);
int find_launcher_p3_nch(int la, int lb, int lc, int ld);
int find_launcher_p4_nch(int la, int lb, int lc, int ld);
int find_launcher_p5_nch(int la, int lb, int lc, int ld);
int find_launcher_p6_nch(int la, int lb, int lc, int ld);

// Wave items per wave unit of a class's plan (its component chunks, or its C
// chunks when the plan carries J[c,d] across A chunks); 0 without a plan.
int eri_fast_nch(int la, int lb, int lc, int ld) {
  int (*finders[])(int, int, int, int) = {find_launcher_p0_nch, find_launcher_p1_nch, find_launcher_p2_nch,
                                          find_launcher_p3_nch, find_launcher_p4_nch, find_launcher_p5_nch,
                                          find_launcher_p6_nch};
  for (auto fn : finders) {
    const int n = fn(la, lb, lc, ld);
    if (n > 0) return n;
  }
  return 0;
}

static asuka_fast::Launcher find_fast(int la, int lb, int lc, int ld) {
  asuka_fast::Launcher f = nullptr;
  asuka_fast::Launcher (*finders[])(int, int, int, int) = {find_launcher_p0, find_launcher_p1, find_launcher_p2, find_launcher_p3, find_launcher_p4, find_launcher_p5, find_launcher_p6};
  for (auto fn : finders)
    if (!f) f = fn(la, lb, lc, ld);
  return f;
}

bool eri_fast_has_plan(int la, int lb, int lc, int ld) { return find_fast(la, lb, lc, ld) != nullptr; }

hipError_t launch_eri_fast(const DevBasis& b, const ClassWork& w, const EriOut& o, hipStream_t s, bool* handled) {
  *handled = false;
  if (w.q_ab || !w.units || (o.mode != MODE_JK && o.mode != MODE_FOCK)) return hipSuccess;
  asuka_fast::Launcher f = find_fast(w.la, w.lb, w.lc, w.ld);
  if (!f) return hipSuccess;
  *handled = true;
  if (w.nunits <= 0) return hipSuccess;
  return f(b, w, o, s);
}
