/*
 * CPU oracle for the integral-direct four-center J/K path.
 *
 * TEST INFRASTRUCTURE ONLY. This file is the checker that the HIP product path
 * is compared against (and the timed CPU baseline). Nothing in the product
 * package links, loads or calls it.
 *
 * It restates the reference's FP64 CPU Rys algorithm (asuka/cueri/_eri_rys_cpu.pyx):
 *   Boys F_m list ................ pyx:187-217   (oracle_boys_list)
 *   roots/weights n=1,2,3 ........ pyx:219-420   (rys_1/2/3)
 *   Stieltjes + QL, n>=4 ......... pyx:422-684   (rys_generic_*)
 *   G recursion .................. pyx:687-723   (compute_G)
 *   binomial HRR shift ........... pyx:818-835   (shift_from_G)
 *   pair-table tile kernel ....... pyx:1096-1291 (tile_sp)
 *   pair tables .................. pyx:2066-2129 (oracle_pair_tables)
 *   Schwarz diag max ............. pyx:2486-2611 (oracle_schwarz)
 * and the J/K digestion semantics of
 *   cueri/cuda/ext/src/cueri_cuda_kernels_dense_scatter.cu:246-352 (oracle_direct_jk)
 * and the dense eri_mat[pq,rs] builder feeding hf/dense_jk.py:55-81 (oracle_dense_eri).
 *
 * Precision. The reference computes roots from ordinary moments in double
 * precision; that map is ill-conditioned (relative error ~2e-13 at n = 3,
 * ~2e-11 at n = 5 and ~5e-9 at n = 7, measured against 40-digit values). With
 * hp != 0 the moment -> Stieltjes -> QL algorithm (the generic path, which for
 * n = 2, 3 is algebraically the same as the reference's closed forms) runs in
 * binary128 (__float128) and is rounded to double, i.e. the reference algorithm
 * in its exact-arithmetic limit;
 * with hp == 0 it runs in double exactly as the reference does (used to measure
 * the reference's own error and as the CPU baseline, which times the
 * reference's arithmetic).
 */
#include <math.h>
#include <quadmath.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define LMAX 4
#define STRIDE (2 * LMAX + 2)
#define GSIZE (STRIDE * STRIDE)
#define NCART_MAX ((LMAX + 1) * (LMAX + 2) / 2)
#define NROOTS_MAX 12

static const double kPi = 3.141592653589793238462643383279502884;
/* 2 * pi^(5/2) */
static const double kTwoPi52 = 2.0 * 3.141592653589793238462643383279502884 * 3.141592653589793238462643383279502884 *
                               1.772453850905516027298167483341145182;

typedef struct {
  int nshell;
  const double* shell_cxyz;      /* (nshell,3) */
  const int32_t* shell_prim_start;
  const int32_t* shell_nprim;
  const int32_t* shell_l;
  const int32_t* shell_ao_start;
  const double* prim_exp;
  const double* prim_coef;
  int nsp;
  const int32_t* sp_A;
  const int32_t* sp_B;
  const int32_t* sp_pair_start;  /* (nsp+1) */
  const double* pair_eta;
  const double* pair_Px;
  const double* pair_Py;
  const double* pair_Pz;
  const double* pair_cK;
} oracle_basis;

static inline int ncart(int l) { return (l + 1) * (l + 2) / 2; }

static double binom(int n, int k) {
  if (k < 0 || k > n) return 0.0;
  if (k == 0 || k == n) return 1.0;
  int kk = k > n - k ? n - k : k;
  double out = 1.0;
  for (int i = 1; i <= kk; ++i) out *= (double)(n - kk + i) / (double)i;
  return out;
}

static void cart_comp(int l, int* lx, int* ly, int* lz) {
  int idx = 0;
  for (int x = l; x >= 0; --x)
    for (int y = l - x; y >= 0; --y) {
      lx[idx] = x;
      ly[idx] = y;
      lz[idx] = l - x - y;
      ++idx;
    }
}

/* ---------------- Boys function (pyx:187-217) ---------------- */

static double boys_f0(double T) {
  if (T < 1e-12) return 1.0 - T / 3.0 + T * T / 10.0;
  return 0.5 * sqrt(kPi / T) * erf(sqrt(T));
}

void oracle_boys_list(double T, int m_max, double* out) {
  if (m_max <= 0) {
    out[0] = boys_f0(T);
    return;
  }
  if (T < 5.0) {
    double term = 1.0, Fm = 0.0;
    for (int k = 0; k < 120; ++k) {
      Fm += term / (double)(2 * m_max + 2 * k + 1);
      term *= -T / (double)(k + 1);
    }
    out[m_max] = Fm;
    double e = exp(-T);
    for (int m = m_max; m > 0; --m) out[m - 1] = (2.0 * T * out[m] + e) / (double)(2 * m - 1);
    return;
  }
  out[0] = boys_f0(T);
  double e = exp(-T);
  for (int m = 1; m <= m_max; ++m) out[m] = ((double)(2 * m - 1) * out[m - 1] - e) / (2.0 * T);
}

static void boys_list_q(__float128 T, int m_max, __float128* out) {
  const __float128 pi = M_PIq;
  if (T < 5.0Q) {
    __float128 term = 1.0Q, Fm = 0.0Q;
    for (int k = 0; k < 200; ++k) {
      Fm += term / (__float128)(2 * m_max + 2 * k + 1);
      term *= -T / (__float128)(k + 1);
    }
    out[m_max] = Fm;
    __float128 e = expq(-T);
    for (int m = m_max; m > 0; --m) out[m - 1] = (2.0Q * T * out[m] + e) / (__float128)(2 * m - 1);
    return;
  }
  out[0] = 0.5Q * sqrtq(pi / T) * erfq(sqrtq(T));
  __float128 e = expq(-T);
  for (int m = 1; m <= m_max; ++m) out[m] = ((__float128)(2 * m - 1) * out[m - 1] - e) / (2.0Q * T);
}

static inline double clamp01(double x) { return x < 0.0 ? 0.0 : (x > 1.0 ? 1.0 : x); }

/* ---------------- closed forms n = 1, 2, 3 (pyx:219-420) ---------------- */

static void rys_1(double T, double* r, double* w) {
  double F[2];
  oracle_boys_list(T, 1, F);
  w[0] = F[0];
  r[0] = clamp01(F[0] > 0.0 ? F[1] / F[0] : 0.0);
}

static void rys_2(double T, double* r, double* w) {
  double mu[4];
  oracle_boys_list(T, 3, mu);
  double a0 = mu[0] > 0.0 ? mu[1] / mu[0] : 0.0;
  double h1 = mu[2] - mu[1] * mu[1] / (mu[0] > 1e-300 ? mu[0] : 1e-300);
  if (h1 < 0.0) h1 = 0.0;
  double b1 = mu[0] > 0.0 ? h1 / mu[0] : 0.0;
  double a1n = mu[3] - 2.0 * a0 * mu[2] + a0 * a0 * mu[1];
  double a1 = h1 > 0.0 ? a1n / h1 : a0;
  double s1 = b1 > 0.0 ? sqrt(b1) : 0.0;
  double tr = a0 + a1, det = a0 * a1 - b1;
  double disc = tr * tr - 4.0 * det;
  if (disc < 0.0) disc = 0.0;
  double rt = sqrt(disc);
  double l0 = 0.5 * (tr - rt), l1 = 0.5 * (tr + rt);
  if (l0 > l1) {
    double t = l0;
    l0 = l1;
    l1 = t;
  }
  r[0] = clamp01(l0);
  r[1] = clamp01(l1);
  double v1 = l0 - a0, n0 = s1 * s1 + v1 * v1;
  w[0] = n0 > 0.0 ? mu[0] * s1 * s1 / n0 : 0.0;
  v1 = l1 - a0;
  n0 = s1 * s1 + v1 * v1;
  w[1] = n0 > 0.0 ? mu[0] * s1 * s1 / n0 : 0.0;
}

static void rys_3(double T, double* r, double* w) {
  double mu[6];
  oracle_boys_list(T, 5, mu);
  double a0 = mu[0] > 0.0 ? mu[1] / mu[0] : 0.0;
  double h1 = mu[2] - mu[1] * mu[1] / (mu[0] > 1e-300 ? mu[0] : 1e-300);
  if (h1 < 0.0) h1 = 0.0;
  double b1 = mu[0] > 0.0 ? h1 / mu[0] : 0.0;
  double a1n = mu[3] - 2.0 * a0 * mu[2] + a0 * a0 * mu[1];
  double a1 = h1 > 0.0 ? a1n / h1 : a0;
  /* P2(x) = x^2 + c1 x + c0 */
  double c1 = -(a0 + a1), c0 = a0 * a1 - b1;
  double h2 = mu[4] + 2.0 * c1 * mu[3] + 2.0 * c0 * mu[2] + c1 * c1 * mu[2] + 2.0 * c1 * c0 * mu[1] + c0 * c0 * mu[0];
  if (h2 < 0.0) h2 = 0.0;
  double b2 = h1 > 0.0 ? h2 / h1 : 0.0;
  double a2n = mu[5] + 2.0 * c1 * mu[4] + 2.0 * c0 * mu[3] + c1 * c1 * mu[3] + 2.0 * c1 * c0 * mu[2] + c0 * c0 * mu[1];
  double a2 = h2 > 0.0 ? a2n / h2 : a1;
  double s1 = b1 > 0.0 ? sqrt(b1) : 0.0, s2 = b2 > 0.0 ? sqrt(b2) : 0.0;
  /* eigenvalues of the symmetric tridiagonal [[a0,s1,0],[s1,a1,s2],[0,s2,a2]] (trigonometric form) */
  double A11 = a0, A22 = a1, A33 = a2, A12 = s1, A13 = 0.0, A23 = s2;
  double p1 = A12 * A12 + A13 * A13 + A23 * A23;
  double l0 = A11, l1 = A22, l2 = A33;
  if (p1 > 0.0) {
    double q = (A11 + A22 + A33) / 3.0;
    double p2 = (A11 - q) * (A11 - q) + (A22 - q) * (A22 - q) + (A33 - q) * (A33 - q) + 2.0 * p1;
    double p = p2 > 0.0 ? sqrt(p2 / 6.0) : 0.0;
    if (p > 0.0) {
      double B11 = (A11 - q) / p, B22 = (A22 - q) / p, B33 = (A33 - q) / p;
      double B12 = A12 / p, B13 = A13 / p, B23 = A23 / p;
      double detB = B11 * (B22 * B33 - B23 * B23) - B12 * (B12 * B33 - B23 * B13) + B13 * (B12 * B23 - B22 * B13);
      double rr = detB * 0.5;
      if (rr < -1.0) rr = -1.0;
      if (rr > 1.0) rr = 1.0;
      double phi = acos(rr) / 3.0;
      l2 = q + 2.0 * p * cos(phi);
      l0 = q + 2.0 * p * cos(phi + 2.0 * kPi / 3.0);
      l1 = 3.0 * q - l0 - l2;
    }
  }
  double e[3] = {l0, l1, l2}, t;
  if (e[0] > e[1]) { t = e[0]; e[0] = e[1]; e[1] = t; }
  if (e[1] > e[2]) { t = e[1]; e[1] = e[2]; e[2] = t; }
  if (e[0] > e[1]) { t = e[0]; e[0] = e[1]; e[1] = t; }
  for (int i = 0; i < 3; ++i) {
    r[i] = clamp01(e[i]);
    double r0x = a0 - e[i], r2z = a2 - e[i];
    double v0 = s1 * r2z, v1 = -r0x * r2z, v2 = r0x * s2;
    double n2 = v0 * v0 + v1 * v1 + v2 * v2;
    w[i] = n2 > 0.0 ? mu[0] * v0 * v0 / n2 : 0.0;
  }
}

/* ---------------- generic Stieltjes + QL (pyx:422-684) ----------------
 * Written once as a macro over the real type so the double and binary128
 * instances are the same algorithm. */

#define DEFINE_RYS_GENERIC(NAME, REAL, SQRT, FABS, COPYSIGN, BOYS, EPS)                              \
  static void NAME(int n, REAL T, REAL* r, REAL* w) {                                                \
    REAL mu[2 * NROOTS_MAX], alpha[NROOTS_MAX], beta[NROOTS_MAX];                                   \
    REAL pp[NROOTS_MAX + 1], pc[NROOTS_MAX + 1], pn[NROOTS_MAX + 1];                                \
    REAL d[NROOTS_MAX], e[NROOTS_MAX], z0[NROOTS_MAX];                                              \
    BOYS(T, 2 * n - 1, mu);                                                                          \
    REAL mu0 = mu[0];                                                                                \
    if (!(mu0 > 0)) {                                                                                \
      for (int i = 0; i < n; ++i) r[i] = w[i] = 0;                                                   \
      return;                                                                                        \
    }                                                                                                \
    for (int i = 0; i <= n; ++i) pp[i] = pc[i] = pn[i] = 0;                                         \
    pc[0] = 1;                                                                                       \
    REAL hc = mu0;                                                                                   \
    alpha[0] = mu[1] / mu0;                                                                          \
    beta[0] = 0;                                                                                     \
    for (int k = 0; k < n - 1; ++k) {                                                                \
      REAL bk = (k >= 1 && hc > 0) ? beta[k] : (REAL)0;                                              \
      for (int i = 0; i <= n; ++i) pn[i] = 0;                                                        \
      for (int i = 0; i <= k; ++i) {                                                                 \
        pn[i + 1] += pc[i];                                                                          \
        pn[i] -= alpha[k] * pc[i];                                                                   \
      }                                                                                              \
      if (k >= 1)                                                                                    \
        for (int i = 0; i < k; ++i) pn[i] -= bk * pp[i];                                             \
      REAL hn = 0, an = 0;                                                                           \
      for (int i = 0; i <= k + 1; ++i) {                                                             \
        if (pn[i] == 0) continue;                                                                    \
        for (int j = 0; j <= k + 1; ++j) {                                                           \
          hn += pn[i] * pn[j] * mu[i + j];                                                           \
          an += pn[i] * pn[j] * mu[i + j + 1];                                                       \
        }                                                                                            \
      }                                                                                              \
      if (hn < 0) hn = 0;                                                                            \
      beta[k + 1] = hc > 0 ? hn / hc : (REAL)0;                                                      \
      alpha[k + 1] = hn > 0 ? an / hn : alpha[k];                                                    \
      for (int i = 0; i <= n; ++i) {                                                                 \
        pp[i] = pc[i];                                                                               \
        pc[i] = pn[i];                                                                               \
      }                                                                                              \
      hc = hn;                                                                                       \
    }                                                                                                \
    for (int i = 0; i < n; ++i) {                                                                    \
      d[i] = alpha[i];                                                                               \
      e[i] = 0;                                                                                      \
    }                                                                                                \
    for (int i = 0; i < n - 1; ++i) e[i] = beta[i + 1] > 0 ? SQRT(beta[i + 1]) : (REAL)0;            \
    /* implicit QL on the Jacobi matrix, tracking the first eigenvector row */                      \
    e[n - 1] = 0;                                                                                    \
    for (int i = 0; i < n; ++i) z0[i] = (i == 0) ? 1 : 0;                                           \
    for (int l = 0; l < n; ++l) {                                                                    \
      int iter = 0;                                                                                  \
      for (;;) {                                                                                     \
        int m;                                                                                       \
        for (m = l; m < n - 1; ++m) {                                                                \
          REAL dd = FABS(d[m]) + FABS(d[m + 1]);                                                     \
          if (dd == 0) dd = 1;                                                                       \
          if (FABS(e[m]) <= EPS * dd) break;                                                         \
        }                                                                                            \
        if (m == l || iter >= 128) break;                                                            \
        ++iter;                                                                                      \
        REAL g = (d[l + 1] - d[l]) / (2 * e[l]);                                                     \
        REAL rr = SQRT(g * g + 1);                                                                   \
        g = d[m] - d[l] + e[l] / (g + COPYSIGN(rr, g));                                              \
        REAL s = 1, c = 1, p = 0;                                                                    \
        int broke = 0;                                                                               \
        for (int i = m - 1; i >= l; --i) {                                                           \
          REAL f = s * e[i], b = c * e[i];                                                           \
          rr = SQRT(f * f + g * g);                                                                  \
          e[i + 1] = rr;                                                                             \
          if (rr == 0) {                                                                             \
            d[i + 1] -= p;                                                                           \
            e[m] = 0;                                                                                \
            broke = 1;                                                                               \
            break;                                                                                   \
          }                                                                                          \
          s = f / rr;                                                                                \
          c = g / rr;                                                                                \
          g = d[i + 1] - p;                                                                          \
          REAL t2 = (d[i] - g) * s + 2 * c * b;                                                      \
          p = s * t2;                                                                                \
          d[i + 1] = g + p;                                                                          \
          g = c * t2 - b;                                                                            \
          REAL z1 = z0[i + 1], zi = z0[i];                                                           \
          z0[i + 1] = s * zi + c * z1;                                                               \
          z0[i] = c * zi - s * z1;                                                                   \
        }                                                                                            \
        if (broke) continue;                                                                         \
        d[l] -= p;                                                                                   \
        e[l] = g;                                                                                    \
        e[m] = 0;                                                                                    \
      }                                                                                              \
    }                                                                                                \
    for (int i = 0; i < n - 1; ++i) { /* ascending sort, carrying z0 */                             \
      int k = i;                                                                                     \
      REAL p = d[i];                                                                                 \
      for (int j = i + 1; j < n; ++j)                                                                \
        if (d[j] < p) {                                                                              \
          k = j;                                                                                     \
          p = d[j];                                                                                  \
        }                                                                                            \
      if (k != i) {                                                                                  \
        REAL t = d[i];                                                                               \
        d[i] = d[k];                                                                                 \
        d[k] = t;                                                                                    \
        t = z0[i];                                                                                   \
        z0[i] = z0[k];                                                                               \
        z0[k] = t;                                                                                   \
      }                                                                                              \
    }                                                                                                \
    for (int i = 0; i < n; ++i) {                                                                    \
      r[i] = d[i] < 0 ? (REAL)0 : (d[i] > 1 ? (REAL)1 : d[i]);                                       \
      w[i] = mu0 * z0[i] * z0[i];                                                                    \
      if (w[i] < 0) w[i] = 0;                                                                        \
    }                                                                                                \
  }

DEFINE_RYS_GENERIC(rys_generic_d, double, sqrt, fabs, copysign, oracle_boys_list, 1e-15)
DEFINE_RYS_GENERIC(rys_generic_q, __float128, sqrtq, fabsq, copysignq, boys_list_q, 1e-32Q)

/* Rys roots (x = t^2) and weights. hp selects binary128 for the generic path. */
void oracle_rys(int n, double T, int hp, double* r, double* w) {
  if (hp && n >= 2) {
    __float128 rq[NROOTS_MAX], wq[NROOTS_MAX];
    rys_generic_q(n, (__float128)T, rq, wq);
    for (int i = 0; i < n; ++i) {
      r[i] = (double)rq[i];
      w[i] = (double)wq[i];
    }
    return;
  }
  if (n == 1) {
    rys_1(T, r, w);
    return;
  }
  if (n == 2) {
    rys_2(T, r, w);
    return;
  }
  if (n == 3) {
    rys_3(T, r, w);
    return;
  }
  rys_generic_d(n, T, r, w);
}

/* ---------------- G recursion and shift (pyx:687-723, 818-835) ---------------- */

static void compute_G(double* G, int nmax, int mmax, double C, double Cp, double B0, double B1, double B1p) {
  G[0] = 1.0;
  if (nmax > 0) G[1 * STRIDE] = C;
  if (mmax > 0) G[1] = Cp;
  for (int a = 2; a <= nmax; ++a) G[a * STRIDE] = B1 * (double)(a - 1) * G[(a - 2) * STRIDE] + C * G[(a - 1) * STRIDE];
  for (int b = 2; b <= mmax; ++b) G[b] = B1p * (double)(b - 1) * G[b - 2] + Cp * G[b - 1];
  if (mmax == 0 || nmax == 0) return;
  for (int a = 1; a <= nmax; ++a) {
    G[a * STRIDE + 1] = (double)a * B0 * G[(a - 1) * STRIDE] + Cp * G[a * STRIDE];
    for (int b = 2; b <= mmax; ++b)
      G[a * STRIDE + b] = B1p * (double)(b - 1) * G[a * STRIDE + b - 2] + (double)a * B0 * G[(a - 1) * STRIDE + b - 1] +
                          Cp * G[a * STRIDE + b - 1];
  }
}

static double shift_from_G(const double* G, int i, int j, int k, int l, const double* xij, const double* xkl) {
  double out = 0.0;
  for (int m = 0; m <= l; ++m) {
    double acc = 0.0;
    for (int n = 0; n <= j; ++n) acc += binom(j, n) * xij[j - n] * G[(n + i) * STRIDE + (m + k)];
    out += binom(l, m) * xkl[l - m] * acc;
  }
  return out;
}

/* ---------------- tile kernel over pair tables (pyx:1096-1291) ----------------
 * out[(ia*nB+ib)*nCD + ic*nD+id] += (ab|cd), accumulated (caller zeroes). */
static void tile_sp(const oracle_basis* bs, int spAB, int spCD, int hp, double* out) {
  int A = bs->sp_A[spAB], B = bs->sp_B[spAB], C = bs->sp_A[spCD], D = bs->sp_B[spCD];
  int la = bs->shell_l[A], lb = bs->shell_l[B], lc = bs->shell_l[C], ld = bs->shell_l[D];
  int nA = ncart(la), nB = ncart(lb), nC = ncart(lc), nD = ncart(ld), nCD = nC * nD;
  const double* cA = bs->shell_cxyz + 3 * A;
  const double* cB = bs->shell_cxyz + 3 * B;
  const double* cC = bs->shell_cxyz + 3 * C;
  const double* cD = bs->shell_cxyz + 3 * D;
  double xij[3][LMAX + 2], xkl[3][LMAX + 2];
  for (int ax = 0; ax < 3; ++ax) {
    xij[ax][0] = 1.0;
    xkl[ax][0] = 1.0;
    for (int p = 1; p <= LMAX + 1; ++p) {
      xij[ax][p] = xij[ax][p - 1] * (cA[ax] - cB[ax]);
      xkl[ax][p] = xkl[ax][p - 1] * (cC[ax] - cD[ax]);
    }
  }
  int Ax_[NCART_MAX], Ay_[NCART_MAX], Az_[NCART_MAX], Bx_[NCART_MAX], By_[NCART_MAX], Bz_[NCART_MAX];
  int Cx_[NCART_MAX], Cy_[NCART_MAX], Cz_[NCART_MAX], Dx_[NCART_MAX], Dy_[NCART_MAX], Dz_[NCART_MAX];
  cart_comp(la, Ax_, Ay_, Az_);
  cart_comp(lb, Bx_, By_, Bz_);
  cart_comp(lc, Cx_, Cy_, Cz_);
  cart_comp(ld, Dx_, Dy_, Dz_);
  int nmax = la + lb, mmax = lc + ld, nroots = (la + lb + lc + ld) / 2 + 1;
  int baseAB = bs->sp_pair_start[spAB], baseCD = bs->sp_pair_start[spCD];
  int nAB_p = bs->sp_pair_start[spAB + 1] - baseAB, nCD_p = bs->sp_pair_start[spCD + 1] - baseCD;
  double roots[NROOTS_MAX], weights[NROOTS_MAX];
  double G[3][GSIZE];
  for (int iab = 0; iab < nAB_p; ++iab) {
    double p = bs->pair_eta[baseAB + iab];
    double P[3] = {bs->pair_Px[baseAB + iab], bs->pair_Py[baseAB + iab], bs->pair_Pz[baseAB + iab]};
    double cKab = bs->pair_cK[baseAB + iab];
    for (int icd = 0; icd < nCD_p; ++icd) {
      double q = bs->pair_eta[baseCD + icd];
      double Q[3] = {bs->pair_Px[baseCD + icd], bs->pair_Py[baseCD + icd], bs->pair_Pz[baseCD + icd]};
      double cKcd = bs->pair_cK[baseCD + icd];
      double denom = p + q, inv_denom = 1.0 / denom;
      double PQ2 = 0.0;
      for (int ax = 0; ax < 3; ++ax) PQ2 += (P[ax] - Q[ax]) * (P[ax] - Q[ax]);
      double T = p * q * inv_denom * PQ2;
      double base = kTwoPi52 / (p * q * sqrt(denom)) * cKab * cKcd;
      oracle_rys(nroots, T, hp, roots, weights);
      for (int u = 0; u < nroots; ++u) {
        double x = roots[u];
        double B0 = x * 0.5 * inv_denom;
        double B1 = (1.0 - x) * 0.5 / p + B0;
        double B1p = (1.0 - x) * 0.5 / q + B0;
        for (int ax = 0; ax < 3; ++ax) {
          double Cc = (P[ax] - cA[ax]) + (q * inv_denom) * x * (Q[ax] - P[ax]);
          double Cp = (Q[ax] - cC[ax]) + (p * inv_denom) * x * (P[ax] - Q[ax]);
          compute_G(G[ax], nmax, mmax, Cc, Cp, B0, B1, B1p);
        }
        double scale = base * weights[u];
        for (int ia = 0; ia < nA; ++ia)
          for (int ib = 0; ib < nB; ++ib) {
            int row = ia * nB + ib;
            for (int ic = 0; ic < nC; ++ic)
              for (int id = 0; id < nD; ++id) {
                double Ix = shift_from_G(G[0], Ax_[ia], Bx_[ib], Cx_[ic], Dx_[id], xij[0], xkl[0]);
                double Iy = shift_from_G(G[1], Ay_[ia], By_[ib], Cy_[ic], Dy_[id], xij[1], xkl[1]);
                double Iz = shift_from_G(G[2], Az_[ia], Bz_[ib], Cz_[ic], Dz_[id], xij[2], xkl[2]);
                out[row * nCD + ic * nD + id] += scale * (Ix * Iy * Iz);
              }
          }
      }
    }
  }
}

/* ---------------- public entry points ---------------- */

int oracle_lmax(void) { return LMAX; }

/* Pair tables (pyx:2066-2129): eta, P, cK = cA cB exp(-ab/eta |AB|^2). */
void oracle_pair_tables(const oracle_basis* bs, double* eta, double* Px, double* Py, double* Pz, double* cK) {
  for (int sp = 0; sp < bs->nsp; ++sp) {
    int A = bs->sp_A[sp], B = bs->sp_B[sp];
    int base = bs->sp_pair_start[sp];
    int nprimB = bs->shell_nprim[B];
    int npair = bs->sp_pair_start[sp + 1] - base;
    const double* a3 = bs->shell_cxyz + 3 * A;
    const double* b3 = bs->shell_cxyz + 3 * B;
    double dx = a3[0] - b3[0], dy = a3[1] - b3[1], dz = a3[2] - b3[2];
    double AB2 = dx * dx + dy * dy + dz * dz;
    for (int idx = 0; idx < npair; ++idx) {
      int ia = idx / nprimB, ib = idx - ia * nprimB;
      int pA = bs->shell_prim_start[A] + ia, pB = bs->shell_prim_start[B] + ib;
      double al = bs->prim_exp[pA], be = bs->prim_exp[pB];
      double e = al + be, ie = 1.0 / e;
      eta[base + idx] = e;
      Px[base + idx] = (al * a3[0] + be * b3[0]) * ie;
      Py[base + idx] = (al * a3[1] + be * b3[1]) * ie;
      Pz[base + idx] = (al * a3[2] + be * b3[2]) * ie;
      cK[base + idx] = bs->prim_coef[pA] * bs->prim_coef[pB] * exp(-(al * be) * ie * AB2);
    }
  }
}

int oracle_eri_tile(const oracle_basis* bs, int spAB, int spCD, int hp, double* out) {
  int la = bs->shell_l[bs->sp_A[spAB]], lb = bs->shell_l[bs->sp_B[spAB]];
  int lc = bs->shell_l[bs->sp_A[spCD]], ld = bs->shell_l[bs->sp_B[spCD]];
  if (la > LMAX || lb > LMAX || lc > LMAX || ld > LMAX) return -1;
  memset(out, 0, sizeof(double) * ncart(la) * ncart(lb) * ncart(lc) * ncart(ld));
  tile_sp(bs, spAB, spCD, hp, out);
  return 0;
}

/* Schwarz bounds Q_AB = sqrt(max_{mu in A, nu in B} (mu nu|mu nu)) (pyx:2486-2611). */
void oracle_schwarz(const oracle_basis* bs, int hp, int nthreads, double* Q) {
#ifdef _OPENMP
  if (nthreads > 0) omp_set_num_threads(nthreads);
#pragma omp parallel
#endif
  {
    double* tile = (double*)malloc(sizeof(double) * NCART_MAX * NCART_MAX * NCART_MAX * NCART_MAX);
#ifdef _OPENMP
#pragma omp for schedule(dynamic, 16)
#endif
    for (int sp = 0; sp < bs->nsp; ++sp) {
      int la = bs->shell_l[bs->sp_A[sp]], lb = bs->shell_l[bs->sp_B[sp]];
      int nAB = ncart(la) * ncart(lb);
      memset(tile, 0, sizeof(double) * nAB * nAB);
      tile_sp(bs, sp, sp, hp, tile);
      double mx = 0.0;
      for (int i = 0; i < nAB; ++i)
        if (tile[i * nAB + i] > mx) mx = tile[i * nAB + i];
      Q[sp] = mx > 0.0 ? sqrt(mx) : 0.0;
    }
    free(tile);
  }
}

/* Digest one tile with the ordered-contraction semantics of
 * dense_scatter.cu:246-352 (J and K accumulated unsymmetrised). */
static void digest(const oracle_basis* bs, int spab, int spcd, const double* tile, int nao, const double* Dm, double* J,
                   double* K) {
  int A = bs->sp_A[spab], B = bs->sp_B[spab], C = bs->sp_A[spcd], Dsh = bs->sp_B[spcd];
  int nA = ncart(bs->shell_l[A]), nB = ncart(bs->shell_l[B]), nC = ncart(bs->shell_l[C]), nD = ncart(bs->shell_l[Dsh]);
  int a0 = bs->shell_ao_start[A], b0 = bs->shell_ao_start[B], c0 = bs->shell_ao_start[C], d0 = bs->shell_ao_start[Dsh];
  int ab_neq = A != B, cd_neq = C != Dsh, bk = spab != spcd;
  long N = nao;
  for (int ia = 0; ia < nA; ++ia)
    for (int ib = 0; ib < nB; ++ib)
      for (int ic = 0; ic < nC; ++ic)
        for (int id = 0; id < nD; ++id) {
          double v = tile[((ia * nB + ib) * nC + ic) * nD + id];
          if (v == 0.0) continue;
          long a = a0 + ia, b = b0 + ib, c = c0 + ic, d = d0 + id;
          if (J) {
            double vab = (cd_neq ? 2.0 : 1.0) * v * Dm[c * N + d];
            J[a * N + b] += vab;
            if (ab_neq) J[b * N + a] += vab;
            if (bk) {
              double vcd = (ab_neq ? 2.0 : 1.0) * v * Dm[a * N + b];
              J[c * N + d] += vcd;
              if (cd_neq) J[d * N + c] += vcd;
            }
          }
          if (K) {
            double t = v * Dm[b * N + d];
            K[a * N + c] += t;
            if (bk) K[c * N + a] += t;
            if (cd_neq) {
              t = v * Dm[b * N + c];
              K[a * N + d] += t;
              if (bk) K[d * N + a] += t;
            }
            if (ab_neq) {
              t = v * Dm[a * N + d];
              K[b * N + c] += t;
              if (bk) K[c * N + b] += t;
              if (cd_neq) {
                t = v * Dm[a * N + c];
                K[b * N + d] += t;
                if (bk) K[d * N + b] += t;
              }
            }
          }
        }
}

/* Direct J/K over an explicit canonical quartet list. J, K are overwritten with
 * the unsymmetrised accumulation (caller symmetrises, hf/direct_jk.py:2190-2195).
 * Threads use private accumulators reduced at the end. */
void oracle_direct_jk(const oracle_basis* bs, long ntasks, const int32_t* task_ab, const int32_t* task_cd, int nao,
                      const double* Dm, double* J, double* K, int hp, int nthreads) {
  long nn = (long)nao * nao;
  if (J) memset(J, 0, sizeof(double) * nn);
  if (K) memset(K, 0, sizeof(double) * nn);
#ifdef _OPENMP
  if (nthreads > 0) omp_set_num_threads(nthreads);
#pragma omp parallel
#endif
  {
    double* tile = (double*)malloc(sizeof(double) * NCART_MAX * NCART_MAX * NCART_MAX * NCART_MAX);
    double* Jp = J ? (double*)calloc(nn, sizeof(double)) : NULL;
    double* Kp = K ? (double*)calloc(nn, sizeof(double)) : NULL;
#ifdef _OPENMP
#pragma omp for schedule(dynamic, 64)
#endif
    for (long t = 0; t < ntasks; ++t) {
      int ab = task_ab[t], cd = task_cd[t];
      int la = bs->shell_l[bs->sp_A[ab]], lb = bs->shell_l[bs->sp_B[ab]];
      int lc = bs->shell_l[bs->sp_A[cd]], ld = bs->shell_l[bs->sp_B[cd]];
      memset(tile, 0, sizeof(double) * ncart(la) * ncart(lb) * ncart(lc) * ncart(ld));
      tile_sp(bs, ab, cd, hp, tile);
      digest(bs, ab, cd, tile, nao, Dm, Jp, Kp);
    }
#ifdef _OPENMP
#pragma omp critical
#endif
    {
      if (J)
        for (long i = 0; i < nn; ++i) J[i] += Jp[i];
      if (K)
        for (long i = 0; i < nn; ++i) K[i] += Kp[i];
    }
    free(tile);
    free(Jp);
    free(Kp);
  }
}

/* Dense eri[(p*nao+q)*nao^2 + r*nao+s] over all canonical quartets, expanded by
 * 8-fold symmetry (the eri_mat the reference's dense J/K einsums consume). */
void oracle_dense_eri(const oracle_basis* bs, int nao, int hp, int nthreads, double* eri) {
  long N = nao, N2 = N * N;
#ifdef _OPENMP
  if (nthreads > 0) omp_set_num_threads(nthreads);
#pragma omp parallel
#endif
  {
    double* tile = (double*)malloc(sizeof(double) * NCART_MAX * NCART_MAX * NCART_MAX * NCART_MAX);
#ifdef _OPENMP
#pragma omp for schedule(dynamic, 1)
#endif
    for (int ab = 0; ab < bs->nsp; ++ab) {
      for (int cd = 0; cd <= ab; ++cd) {
        int A = bs->sp_A[ab], B = bs->sp_B[ab], C = bs->sp_A[cd], Dsh = bs->sp_B[cd];
        int nA = ncart(bs->shell_l[A]), nB = ncart(bs->shell_l[B]), nC = ncart(bs->shell_l[C]), nD = ncart(bs->shell_l[Dsh]);
        memset(tile, 0, sizeof(double) * nA * nB * nC * nD);
        tile_sp(bs, ab, cd, hp, tile);
        int a0 = bs->shell_ao_start[A], b0 = bs->shell_ao_start[B], c0 = bs->shell_ao_start[C], d0 = bs->shell_ao_start[Dsh];
        for (int ia = 0; ia < nA; ++ia)
          for (int ib = 0; ib < nB; ++ib)
            for (int ic = 0; ic < nC; ++ic)
              for (int id = 0; id < nD; ++id) {
                double v = tile[((ia * nB + ib) * nC + ic) * nD + id];
                long a = a0 + ia, b = b0 + ib, c = c0 + ic, d = d0 + id;
                long pq[2] = {a * N + b, b * N + a}, rs[2] = {c * N + d, d * N + c};
                for (int i = 0; i < 2; ++i)
                  for (int j = 0; j < 2; ++j) {
                    eri[pq[i] * N2 + rs[j]] = v;
                    eri[rs[j] * N2 + pq[i]] = v;
                  }
              }
      }
    }
    free(tile);
  }
}

/* ---------------- nuclear derivatives (pyx:4316-4656) ----------------
 * out[center*3 + xyz] += sum_{abcd} bar[(ia*nB+ib)*nCD + ic*nD+id] d(ab|cd)/dR_center,xyz
 * for the four centres [A, B, C, D], each from its own primitive exponent:
 * d/dA_x = 2 alpha (a+1_x ..) - a_x (a-1_x ..). Roots floor((L+1)/2)+1, tables to
 * la+lb+1 and lc+ld+1. The pair index decodes to primitives as ia = idx / nprimB. */
static void deriv_contract_sp(const oracle_basis* bs, int spAB, int spCD, int hp, const double* bar, double* out) {
  int A = bs->sp_A[spAB], B = bs->sp_B[spAB], C = bs->sp_A[spCD], D = bs->sp_B[spCD];
  int la = bs->shell_l[A], lb = bs->shell_l[B], lc = bs->shell_l[C], ld = bs->shell_l[D];
  int nA = ncart(la), nB = ncart(lb), nC = ncart(lc), nD = ncart(ld), nCD = nC * nD;
  const double* cA = bs->shell_cxyz + 3 * A;
  const double* cB = bs->shell_cxyz + 3 * B;
  const double* cC = bs->shell_cxyz + 3 * C;
  const double* cD = bs->shell_cxyz + 3 * D;
  double xij[3][LMAX + 3], xkl[3][LMAX + 3];
  for (int ax = 0; ax < 3; ++ax) {
    xij[ax][0] = xkl[ax][0] = 1.0;
    for (int p = 1; p <= LMAX + 2; ++p) {
      xij[ax][p] = xij[ax][p - 1] * (cA[ax] - cB[ax]);
      xkl[ax][p] = xkl[ax][p - 1] * (cC[ax] - cD[ax]);
    }
  }
  int e[4][3][NCART_MAX];  /* Cartesian exponents per centre, axis, component */
  cart_comp(la, e[0][0], e[0][1], e[0][2]);
  cart_comp(lb, e[1][0], e[1][1], e[1][2]);
  cart_comp(lc, e[2][0], e[2][1], e[2][2]);
  cart_comp(ld, e[3][0], e[3][1], e[3][2]);
  int nmax = la + lb + 1, mmax = lc + ld + 1, nroots = (la + lb + lc + ld + 1) / 2 + 1;
  int baseAB = bs->sp_pair_start[spAB], baseCD = bs->sp_pair_start[spCD];
  int nAB_p = bs->sp_pair_start[spAB + 1] - baseAB, nCD_p = bs->sp_pair_start[spCD + 1] - baseCD;
  int nprimB = bs->shell_nprim[B], nprimD = bs->shell_nprim[D];
  double roots[NROOTS_MAX], weights[NROOTS_MAX];
  double G[3][GSIZE];
  for (int iab = 0; iab < nAB_p; ++iab) {
    double p = bs->pair_eta[baseAB + iab];
    double P[3] = {bs->pair_Px[baseAB + iab], bs->pair_Py[baseAB + iab], bs->pair_Pz[baseAB + iab]};
    double cKab = bs->pair_cK[baseAB + iab];
    int ipA = iab / nprimB, ipB = iab - ipA * nprimB;
    double ex[4];
    ex[0] = bs->prim_exp[bs->shell_prim_start[A] + ipA];
    ex[1] = bs->prim_exp[bs->shell_prim_start[B] + ipB];
    for (int icd = 0; icd < nCD_p; ++icd) {
      double q = bs->pair_eta[baseCD + icd];
      double Q[3] = {bs->pair_Px[baseCD + icd], bs->pair_Py[baseCD + icd], bs->pair_Pz[baseCD + icd]};
      double cKcd = bs->pair_cK[baseCD + icd];
      int ipC = icd / nprimD, ipD = icd - ipC * nprimD;
      ex[2] = bs->prim_exp[bs->shell_prim_start[C] + ipC];
      ex[3] = bs->prim_exp[bs->shell_prim_start[D] + ipD];
      double denom = p + q, inv_denom = 1.0 / denom;
      double PQ2 = 0.0;
      for (int ax = 0; ax < 3; ++ax) PQ2 += (P[ax] - Q[ax]) * (P[ax] - Q[ax]);
      double T = p * q * inv_denom * PQ2;
      double base = kTwoPi52 / (p * q * sqrt(denom)) * cKab * cKcd;
      oracle_rys(nroots, T, hp, roots, weights);
      for (int u = 0; u < nroots; ++u) {
        double x = roots[u];
        double B0 = x * 0.5 * inv_denom;
        double B1 = (1.0 - x) * 0.5 / p + B0;
        double B1p = (1.0 - x) * 0.5 / q + B0;
        for (int ax = 0; ax < 3; ++ax) {
          double Cc = (P[ax] - cA[ax]) + (q * inv_denom) * x * (Q[ax] - P[ax]);
          double Cp = (Q[ax] - cC[ax]) + (p * inv_denom) * x * (P[ax] - Q[ax]);
          compute_G(G[ax], nmax, mmax, Cc, Cp, B0, B1, B1p);
        }
        double scale = base * weights[u];
        for (int ia = 0; ia < nA; ++ia)
          for (int ib = 0; ib < nB; ++ib)
            for (int ic = 0; ic < nC; ++ic)
              for (int id = 0; id < nD; ++id) {
                double bv = bar[(ia * nB + ib) * nCD + ic * nD + id];
                if (bv == 0.0) continue;
                int comp[4] = {ia, ib, ic, id};
                double I[3];
                int n4[3][4];
                for (int ax = 0; ax < 3; ++ax) {
                  for (int c = 0; c < 4; ++c) n4[ax][c] = e[c][ax][comp[c]];
                  I[ax] = shift_from_G(G[ax], n4[ax][0], n4[ax][1], n4[ax][2], n4[ax][3], xij[ax], xkl[ax]);
                }
                for (int c = 0; c < 4; ++c)
                  for (int ax = 0; ax < 3; ++ax) {
                    int nn[4] = {n4[ax][0], n4[ax][1], n4[ax][2], n4[ax][3]};
                    nn[c] += 1;
                    double up = shift_from_G(G[ax], nn[0], nn[1], nn[2], nn[3], xij[ax], xkl[ax]);
                    double dn = 0.0;
                    if (n4[ax][c] > 0) {
                      nn[c] -= 2;
                      dn = shift_from_G(G[ax], nn[0], nn[1], nn[2], nn[3], xij[ax], xkl[ax]);
                    }
                    double dI = 2.0 * ex[c] * up - (double)n4[ax][c] * dn;
                    double other = I[(ax + 1) % 3] * I[(ax + 2) % 3];
                    out[c * 3 + ax] += bv * scale * dI * other;
                  }
              }
      }
    }
  }
}

/* Adjoint tile for Tr(DwJ J(Dc)) - 1/2 Tr(DwK K(Dc)) of one ordered pair of
 * shell pairs, in (AB|CD) layout with the pair folding of
 * asuka/mcscf/nac/_dense.py:195-250 (there DwJ = DwK). */
static void bar_tile(const oracle_basis* bs, int spAB, int spCD, int nao, const double* DwJ, const double* DwK,
                     const double* Dc, double* bar) {
  int A = bs->sp_A[spAB], B = bs->sp_B[spAB], C = bs->sp_A[spCD], D = bs->sp_B[spCD];
  int nA = ncart(bs->shell_l[A]), nB = ncart(bs->shell_l[B]), nC = ncart(bs->shell_l[C]), nD = ncart(bs->shell_l[D]);
  long a0 = bs->shell_ao_start[A], b0 = bs->shell_ao_start[B], c0 = bs->shell_ao_start[C], d0 = bs->shell_ao_start[D];
  int same_ab = A == B, same_cd = C == D;
  long N = nao;
  for (int ia = 0; ia < nA; ++ia)
    for (int ib = 0; ib < nB; ++ib)
      for (int ic = 0; ic < nC; ++ic)
        for (int id = 0; id < nD; ++id) {
          long a = a0 + ia, b = b0 + ib, c = c0 + ic, d = d0 + id;
          double v = 0.0;
          if (DwJ) {
            double wab = DwJ[a * N + b] + (same_ab ? 0.0 : DwJ[b * N + a]);
            double dcd = Dc[c * N + d] + (same_cd ? 0.0 : Dc[d * N + c]);
            v += wab * dcd;
          }
          if (DwK) {
            double t = DwK[a * N + c] * Dc[b * N + d];
            if (!same_ab) t += DwK[b * N + c] * Dc[a * N + d];
            if (!same_cd) t += DwK[a * N + d] * Dc[b * N + c];
            if (!same_ab && !same_cd) t += DwK[b * N + d] * Dc[a * N + c];
            v += -0.5 * t;
          }
          bar[(ia * nB + ib) * nC * nD + ic * nD + id] = v;
        }
}

/* Per-quartet derivative contraction with a caller-given adjoint tile. */
int oracle_eri_deriv(const oracle_basis* bs, int spAB, int spCD, int hp, const double* bar, double* out12) {
  memset(out12, 0, sizeof(double) * 12);
  deriv_contract_sp(bs, spAB, spCD, hp, bar, out12);
  return 0;
}

/* Two-electron gradient of Tr(DwJ J(Dc)) - 1/2 Tr(DwK K(Dc)) with the
 * reference's driver loop (asuka/mcscf/nac/_dense.py:1513-1839): every ordered
 * pair (spAB, spCD) of shell pairs, no screening, per-quartet (4, 3) output
 * scattered to the atoms of A, B, C, D. grad (natm, 3) is overwritten. */
void oracle_grad2e(const oracle_basis* bs, int nao, const double* DwJ, const double* DwK, const double* Dc,
                   const int32_t* shell_atom, int natm, int hp, int nthreads, double* grad) {
  memset(grad, 0, sizeof(double) * 3 * natm);
#ifdef _OPENMP
  if (nthreads > 0) omp_set_num_threads(nthreads);
#pragma omp parallel
#endif
  {
    double* bar = (double*)malloc(sizeof(double) * NCART_MAX * NCART_MAX * NCART_MAX * NCART_MAX);
    double* gp = (double*)calloc(3 * (size_t)natm, sizeof(double));
#ifdef _OPENMP
#pragma omp for schedule(dynamic, 1)
#endif
    for (int ab = 0; ab < bs->nsp; ++ab)
      for (int cd = 0; cd < bs->nsp; ++cd) {
        double out[12] = {0};
        bar_tile(bs, ab, cd, nao, DwJ, DwK, Dc, bar);
        deriv_contract_sp(bs, ab, cd, hp, bar, out);
        int at[4] = {shell_atom[bs->sp_A[ab]], shell_atom[bs->sp_B[ab]], shell_atom[bs->sp_A[cd]],
                     shell_atom[bs->sp_B[cd]]};
        for (int c = 0; c < 4; ++c)
          for (int ax = 0; ax < 3; ++ax) gp[3 * at[c] + ax] += out[3 * c + ax];
      }
#ifdef _OPENMP
#pragma omp critical
#endif
    for (int i = 0; i < 3 * natm; ++i) grad[i] += gp[i];
    free(bar);
    free(gp);
  }
}

/* The same contraction over a list of (spAB, spCD) tasks only (a bounded
 * sample of the quartets, for timing): grad (natm, 3) is overwritten with
 * the sum of the tasks' derivative contributions. */
void oracle_grad2e_tasks(const oracle_basis* bs, int nao, const double* DwJ, const double* DwK, const double* Dc,
                         const int32_t* shell_atom, int natm, int hp, int nthreads, long ntask,
                         const int32_t* task_ab, const int32_t* task_cd, double* grad) {
  memset(grad, 0, sizeof(double) * 3 * natm);
#ifdef _OPENMP
  if (nthreads > 0) omp_set_num_threads(nthreads);
#pragma omp parallel
#endif
  {
    double* bar = (double*)malloc(sizeof(double) * NCART_MAX * NCART_MAX * NCART_MAX * NCART_MAX);
    double* gp = (double*)calloc(3 * (size_t)natm, sizeof(double));
#ifdef _OPENMP
#pragma omp for schedule(dynamic, 16)
#endif
    for (long t = 0; t < ntask; ++t) {
      const int ab = task_ab[t], cd = task_cd[t];
      double out[12] = {0};
      bar_tile(bs, ab, cd, nao, DwJ, DwK, Dc, bar);
      deriv_contract_sp(bs, ab, cd, hp, bar, out);
      int at[4] = {shell_atom[bs->sp_A[ab]], shell_atom[bs->sp_B[ab]], shell_atom[bs->sp_A[cd]],
                   shell_atom[bs->sp_B[cd]]};
      for (int c = 0; c < 4; ++c)
        for (int ax = 0; ax < 3; ++ax) gp[3 * at[c] + ax] += out[3 * c + ax];
    }
#ifdef _OPENMP
#pragma omp critical
#endif
    for (int i = 0; i < 3 * natm; ++i) grad[i] += gp[i];
    free(bar);
    free(gp);
  }
}
