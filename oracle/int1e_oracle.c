/*
 * CPU oracle for the one-electron integrals S, T, V (Cartesian AOs).
 *
 * TEST INFRASTRUCTURE ONLY: used by tests and the CPU baseline as the checker.
 *
 * Restates asuka/integrals/int1e_cart.py:
 *   1-D overlap table ............ :77-115  (overlap_1d)
 *   Hermite E coefficients ....... :118-171 (hermite_E)
 *   Coulomb R_tuv^n auxiliaries .. :174-226 (coulomb_R)
 *   kinetic from overlap ......... :758-787
 *   S / T / V builders ........... :379-612, :1031-1190
 * Boys values come from oracle_boys_list (the reference pyx:187-217 form).
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

void oracle_boys_list(double T, int m_max, double* out);

#define L1MAX 5
#define LT (2 * L1MAX + 3)

static inline int nc(int l) { return (l + 1) * (l + 2) / 2; }

static void comps(int l, int* lx, int* ly, int* lz) {
  int idx = 0;
  for (int x = l; x >= 0; --x)
    for (int y = l - x; y >= 0; --y) {
      lx[idx] = x;
      ly[idx] = y;
      lz[idx] = l - x - y;
      ++idx;
    }
}

/* S[i][j], i <= la, j <= lb (int1e_cart.py:77-115) */
static void overlap_1d(int la, int lb, double a, double b, double Ax, double Bx, double S[LT][LT]) {
  double p = a + b, ip = 1.0 / p, mu = a * b * ip;
  double Px = (a * Ax + b * Bx) * ip, PA = Px - Ax, PB = Px - Bx, AB = Ax - Bx;
  memset(S, 0, sizeof(double) * LT * LT);
  S[0][0] = sqrt(M_PI * ip) * exp(-mu * AB * AB);
  double i2p = 0.5 * ip;
  for (int j = 0; j < lb; ++j) {
    S[0][j + 1] = PB * S[0][j];
    if (j > 0) S[0][j + 1] += (double)j * i2p * S[0][j - 1];
  }
  for (int i = 0; i < la; ++i) {
    S[i + 1][0] = PA * S[i][0];
    if (i > 0) S[i + 1][0] += (double)i * i2p * S[i - 1][0];
    for (int j = 0; j < lb; ++j) {
      S[i + 1][j + 1] = PA * S[i][j + 1];
      if (i > 0) S[i + 1][j + 1] += (double)i * i2p * S[i - 1][j + 1];
      S[i + 1][j + 1] += (double)(j + 1) * i2p * S[i][j];
    }
  }
}

static double kin1d(double S[LT][LT], int i, int j, double b) {
  double v = b * (2.0 * j + 1.0) * S[i][j] - 2.0 * b * b * S[i][j + 2];
  if (j >= 2) v -= 0.5 * (double)(j * (j - 1)) * S[i][j - 2];
  return v;
}

/* E[i][j][t] (int1e_cart.py:118-171) */
static void hermite_E(int la, int lb, double a, double b, double Ax, double Bx, double E[L1MAX + 1][L1MAX + 1][2 * L1MAX + 2]) {
  double p = a + b, ip = 1.0 / p, mu = a * b * ip;
  double Px = (a * Ax + b * Bx) * ip, PA = Px - Ax, PB = Px - Bx, AB = Ax - Bx;
  memset(E, 0, sizeof(double) * (L1MAX + 1) * (L1MAX + 1) * (2 * L1MAX + 2));
  E[0][0][0] = exp(-mu * AB * AB);
  double i2p = 0.5 * ip;
  for (int i = 0; i < la; ++i)
    for (int t = 0; t <= i + 1; ++t) {
      double v = PA * E[i][0][t];
      if (t > 0) v += i2p * E[i][0][t - 1];
      if (t + 1 <= i) v += (double)(t + 1) * E[i][0][t + 1];
      E[i + 1][0][t] = v;
    }
  for (int i = 0; i <= la; ++i)
    for (int j = 0; j < lb; ++j)
      for (int t = 0; t <= i + j + 1; ++t) {
        double v = PB * E[i][j][t];
        if (t > 0) v += i2p * E[i][j][t - 1];
        if (t + 1 <= i + j) v += (double)(t + 1) * E[i][j][t + 1];
        E[i][j + 1][t] = v;
      }
}

#define RN (2 * L1MAX + 2)
/* R[n][t][u][v] with n+t+u+v <= nmax (int1e_cart.py:174-226) */
static void coulomb_R(double p, const double PC[3], int nmax, double* R /* RN^4 */) {
#define RI(n, t, u, v) R[(((n) * RN + (t)) * RN + (u)) * RN + (v)]
  double T = p * (PC[0] * PC[0] + PC[1] * PC[1] + PC[2] * PC[2]);
  double F[RN];
  oracle_boys_list(T, nmax, F);
  double fac = -2.0 * p, pw = 1.0;
  for (int n = 0; n <= nmax; ++n) {
    RI(n, 0, 0, 0) = pw * F[n];
    pw *= fac;
  }
  for (int n = nmax - 1; n >= 0; --n) {
    int mm = nmax - n;
    for (int t = 0; t <= mm; ++t)
      for (int u = 0; u <= mm - t; ++u)
        for (int v = 0; v <= mm - t - u; ++v) {
          if (t == 0 && u == 0 && v == 0) continue;
          double val;
          if (t > 0) {
            val = PC[0] * RI(n + 1, t - 1, u, v);
            if (t >= 2) val += (double)(t - 1) * RI(n + 1, t - 2, u, v);
          } else if (u > 0) {
            val = PC[1] * RI(n + 1, t, u - 1, v);
            if (u >= 2) val += (double)(u - 1) * RI(n + 1, t, u - 2, v);
          } else {
            val = PC[2] * RI(n + 1, t, u, v - 1);
            if (v >= 2) val += (double)(v - 1) * RI(n + 1, t, u, v - 2);
          }
          RI(n, t, u, v) = val;
        }
  }
#undef RI
}

/* S, T, V (nao x nao, row-major). Any output pointer may be NULL. */
int oracle_int1e(int nshell, const double* cxyz, const int32_t* pstart, const int32_t* nprim, const int32_t* shl,
                 const int32_t* aostart, const double* pexp, const double* pcoef, int natm, const double* atom_xyz,
                 const double* atom_Z, int nao, double* Sout, double* Tout, double* Vout) {
  double* R = (double*)malloc(sizeof(double) * RN * RN * RN * RN);
  for (int A = 0; A < nshell; ++A) {
    int la = shl[A];
    if (la > L1MAX) {
      free(R);
      return -1;
    }
    for (int B = 0; B <= A; ++B) {
      int lb = shl[B];
      if (lb > L1MAX) {
        free(R);
        return -1;
      }
      int nA = nc(la), nB = nc(lb);
      int ax[21], ay[21], az[21], bx[21], by[21], bz[21];
      comps(la, ax, ay, az);
      comps(lb, bx, by, bz);
      double s[21 * 21], t[21 * 21], v[21 * 21];
      memset(s, 0, sizeof s);
      memset(t, 0, sizeof t);
      memset(v, 0, sizeof v);
      const double* cA = cxyz + 3 * A;
      const double* cB = cxyz + 3 * B;
      for (int ia = 0; ia < nprim[A]; ++ia) {
        double a = pexp[pstart[A] + ia], ca = pcoef[pstart[A] + ia];
        for (int ib = 0; ib < nprim[B]; ++ib) {
          double b = pexp[pstart[B] + ib], cb = pcoef[pstart[B] + ib];
          double c = ca * cb;
          if (c == 0.0) continue;
          double Sx[LT][LT], Sy[LT][LT], Sz[LT][LT];
          overlap_1d(la, lb + 2, a, b, cA[0], cB[0], Sx);
          overlap_1d(la, lb + 2, a, b, cA[1], cB[1], Sy);
          overlap_1d(la, lb + 2, a, b, cA[2], cB[2], Sz);
          for (int i = 0; i < nA; ++i)
            for (int j = 0; j < nB; ++j) {
              double sx = Sx[ax[i]][bx[j]], sy = Sy[ay[i]][by[j]], sz = Sz[az[i]][bz[j]];
              s[i * nB + j] += c * sx * sy * sz;
              double tx = kin1d(Sx, ax[i], bx[j], b), ty = kin1d(Sy, ay[i], by[j], b), tz = kin1d(Sz, az[i], bz[j], b);
              t[i * nB + j] += c * (tx * sy * sz + sx * ty * sz + sx * sy * tz);
            }
          if (Vout) {
            double Ex[L1MAX + 1][L1MAX + 1][2 * L1MAX + 2], Ey[L1MAX + 1][L1MAX + 1][2 * L1MAX + 2],
                Ez[L1MAX + 1][L1MAX + 1][2 * L1MAX + 2];
            hermite_E(la, lb, a, b, cA[0], cB[0], Ex);
            hermite_E(la, lb, a, b, cA[1], cB[1], Ey);
            hermite_E(la, lb, a, b, cA[2], cB[2], Ez);
            double p = a + b;
            double P[3] = {(a * cA[0] + b * cB[0]) / p, (a * cA[1] + b * cB[1]) / p, (a * cA[2] + b * cB[2]) / p};
            double pref = 2.0 * M_PI / p;
            for (int ic = 0; ic < natm; ++ic) {
              double Z = atom_Z[ic];
              if (Z == 0.0) continue;
              double PC[3] = {P[0] - atom_xyz[3 * ic], P[1] - atom_xyz[3 * ic + 1], P[2] - atom_xyz[3 * ic + 2]};
              coulomb_R(p, PC, la + lb, R);
              for (int i = 0; i < nA; ++i)
                for (int j = 0; j < nB; ++j) {
                  double acc = 0.0;
                  for (int tt = 0; tt <= ax[i] + bx[j]; ++tt) {
                    double ex = Ex[ax[i]][bx[j]][tt];
                    if (ex == 0.0) continue;
                    for (int u = 0; u <= ay[i] + by[j]; ++u) {
                      double ey = Ey[ay[i]][by[j]][u];
                      if (ey == 0.0) continue;
                      for (int vv = 0; vv <= az[i] + bz[j]; ++vv) {
                        double ez = Ez[az[i]][bz[j]][vv];
                        if (ez == 0.0) continue;
                        acc += ex * ey * ez * R[((0 * RN + tt) * RN + u) * RN + vv];
                      }
                    }
                  }
                  v[i * nB + j] += c * (-Z) * pref * acc;
                }
            }
          }
        }
      }
      for (int i = 0; i < nA; ++i)
        for (int j = 0; j < nB; ++j) {
          long r = aostart[A] + i, cI = aostart[B] + j;
          if (Sout) Sout[r * nao + cI] = Sout[cI * nao + r] = s[i * nB + j];
          if (Tout) Tout[r * nao + cI] = Tout[cI * nao + r] = t[i * nB + j];
          if (Vout) Vout[r * nao + cI] = Vout[cI * nao + r] = v[i * nB + j];
        }
    }
  }
  free(R);
  return 0;
}
