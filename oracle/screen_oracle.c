/*
 * CPU oracle of the quartet screening of the direct J/K path.
 *
 * TEST INFRASTRUCTURE ONLY: the checker the HIP context's quartet lists and
 * the kernels' per-build counters are compared with. Nothing in the product
 * package links, loads or calls it.
 *
 * Restates, for counting:
 *   canonical Schwarz-screened quartets over the Q-descending order
 *     (asuka/hf/direct_jk.py:1228-1246, asuka/cueri/tasks.py:49-115
 *     build_tasks_screened_sorted_q): rank i pairs with ranks j <= i whose
 *     Q_j >= eps / Q_i (jmax by a binary search on the sorted bounds);
 *   the CSAM density mask Q_ab Q_cd max(D_sp[ab], D_sp[cd]) >= eps_density
 *     (asuka/hf/direct_jk.py:1893-1909), evaluated in that order;
 * and bins every kept quartet as the kernels' counters do, by class pair and
 * symmetry bin (A!=B)<<2 | (C!=D)<<1 | (bra pair != ket pair), with the bra
 * being the pair of the higher pair class and, inside one class, the pair of
 * the later rank (the reference's row index i). Index 8 of a bin row sums the
 * primitive quartets nprim(bra) * nprim(ket).
 */
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define SCREEN_ROW 9

/* number of j in [0, n) with qs[j] >= thr, for qs sorted descending:
 * numpy searchsorted(-qs, -thr, side="right") */
static long count_ge(const double* qs, long n, double thr) {
  long lo = 0, hi = n;
  while (lo < hi) {
    const long mid = lo + (hi - lo) / 2;
    if (-qs[mid] <= -thr)
      lo = mid + 1;
    else
      hi = mid;
  }
  return lo;
}

/*
 * perm: the stable Q-descending order of the shell pairs (numpy
 *   argsort(-Q, kind="stable")), n_valid: how many of them have Q > 0;
 * Q, Dsp (NULL: no density mask), cls (pair class la(la+1)/2 + lb), neq
 *   (A != B), nprim: per shell pair;
 * out: ncls * ncls rows of SCREEN_ROW int64, row X * ncls + Y for the
 *   (bra class X, ket class Y) pair, X >= Y; zeroed here.
 */
void oracle_screen_counts(long nsp, const int64_t* perm, long n_valid, const double* Q, const double* Dsp,
                          double eps_schwarz, double eps_density, const int32_t* cls, const int32_t* neq,
                          const int64_t* nprim, int ncls, int nthreads, int64_t* out) {
  (void)nsp;
  const long nrow = (long)ncls * ncls * SCREEN_ROW;
  memset(out, 0, sizeof(int64_t) * (size_t)nrow);
  double* qs = (double*)malloc(sizeof(double) * (size_t)(n_valid > 0 ? n_valid : 1));
  for (long i = 0; i < n_valid; ++i) qs[i] = Q[perm[i]];
#ifdef _OPENMP
  if (nthreads > 0) omp_set_num_threads(nthreads);
#endif
#pragma omp parallel
  {
    int64_t* loc = (int64_t*)calloc((size_t)nrow, sizeof(int64_t));
#pragma omp for schedule(dynamic, 64)
    for (long i = 0; i < n_valid; ++i) {
      long jmax = i + 1;
      if (eps_schwarz > 0.0) {
        const double q = qs[i] > 1e-300 ? qs[i] : 1e-300;
        const long jm = count_ge(qs, n_valid, eps_schwarz / q);
        if (jm < jmax) jmax = jm;
      }
      const long a = perm[i];
      for (long j = 0; j < jmax; ++j) {
        const long c = perm[j];
        if (Dsp && eps_density > 0.0) {
          const double d = Dsp[a] > Dsp[c] ? Dsp[a] : Dsp[c];
          if (!(Q[a] * Q[c] * d >= eps_density)) continue;
        }
        long bra = a, ket = c;
        if (cls[c] > cls[a]) {
          bra = c;
          ket = a;
        }
        const int bin = (neq[bra] << 2) | (neq[ket] << 1) | (bra != ket);
        int64_t* row = loc + ((long)cls[bra] * ncls + cls[ket]) * SCREEN_ROW;
        row[bin] += 1;
        row[8] += nprim[bra] * nprim[ket];
      }
    }
#pragma omp critical
    for (long k = 0; k < nrow; ++k) out[k] += loc[k];
    free(loc);
  }
  free(qs);
}
