// One-electron part of the nuclear gradient:
//   dL1/dR  with  L1 = Tr(Dh (T + V)) - Tr(Ws S),
// holding the symmetric matrices Dh and Ws fixed (for RHF, Dh = D and Ws the
// energy-weighted density), as one GPU thread per shell pair.
//
// The reference forms the same pieces on the CPU: contract_dhcore_cart
// (asuka/integrals/int1e_cart.py:1738-) for d(T+V) against D, and
// contract_dS_ip_cart (:1545-1737) for the overlap (Pulay) term, assembled in
// asuka/mcscf/nuc_grad_direct.py:151-181.
//
// Centre derivatives of a primitive Cartesian Gaussian,
//   d/dA_x <a|O|b> = 2 alpha <a+1_x|O|b> - a_x <a-1_x|O|b>,
// apply to S, T and each nucleus' V; the derivative with respect to the
// position of nucleus C (operator derivative) follows from translational
// invariance of <a|1/|r-C||b>: dV^C/dC = -(dV^C/dA + dV^C/dB).
// Overlap and kinetic use Obara-Saika 1-D tables, nuclear attraction the Rys
// form of the forward kernel (same conventions as the S/T/V build).
#include "jk_internal.h"
#include "rys_eval.h"

namespace {

constexpr int L1 = ASUKA_LMAX;
constexpr int RI = L1 + 2;  // rows: a up to la + 1
constexpr int RJ = L1 + 4;  // cols: b up to lb + 1, + 2 for the kinetic operator

__device__ void ovl1d(int imax, int jmax, double a, double b, double A, double B, double S[RI][RJ]) {
  const double p = a + b, ip = 1.0 / p, mu = a * b * ip;
  const double Px = (a * A + b * B) * ip, PA = Px - A, PB = Px - B, AB = A - B;
  for (int i = 0; i < RI; ++i)
    for (int j = 0; j < RJ; ++j) S[i][j] = 0.0;
  S[0][0] = sqrt(3.14159265358979323846 * ip) * exp(-mu * AB * AB);
  const double i2p = 0.5 * ip;
  for (int j = 0; j < jmax; ++j) S[0][j + 1] = PB * S[0][j] + (j > 0 ? j * i2p * S[0][j - 1] : 0.0);
  for (int i = 0; i < imax; ++i) {
    S[i + 1][0] = PA * S[i][0] + (i > 0 ? i * i2p * S[i - 1][0] : 0.0);
    for (int j = 0; j < jmax; ++j)
      S[i + 1][j + 1] = PA * S[i][j + 1] + (i > 0 ? i * i2p * S[i - 1][j + 1] : 0.0) + (j + 1) * i2p * S[i][j];
  }
}

// 1-D kinetic factor -1/2 <i| d2/dx2 |j> with the ket exponent b.
__device__ inline double kin1(const double S[RI][RJ], int i, int j, double b) {
  double v = b * (2.0 * j + 1.0) * S[i][j] - 2.0 * b * b * S[i][j + 2];
  if (j >= 2) v -= 0.5 * (double)(j * (j - 1)) * S[i][j - 2];
  return v;
}

// centre derivative of a 1-D factor along its own axis: 2e f(n+1) - n f(n-1)
__device__ inline double dA1(const double S[RI][RJ], int i, int j, double e2) {
  return e2 * S[i + 1][j] - (i > 0 ? i * S[i - 1][j] : 0.0);
}
__device__ inline double dB1(const double S[RI][RJ], int i, int j, double e2) {
  return e2 * S[i][j + 1] - (j > 0 ? j * S[i][j - 1] : 0.0);
}

// 1-D Rys nuclear-attraction table N[i][j] (i <= imax, j <= jmax) for one root.
__device__ void nuc_tab(int imax, int jmax, double PA, double PC, double AB, double x, double i2p,
                        double N[RI][RJ]) {
  double G[2 * L1 + 3];
  const double c0 = PA - x * PC, b1 = (1.0 - x) * i2p;
  const int n = imax + jmax;
  G[0] = 1.0;
  if (n > 0) G[1] = c0;
  for (int a = 1; a < n; ++a) G[a + 1] = c0 * G[a] + a * b1 * G[a - 1];
  for (int i = 0; i <= imax; ++i)
    for (int j = 0; j <= jmax; ++j) {
      if (i + j > n) continue;
      double out = 0.0, binom = 1.0;
      for (int k = 0; k <= j; ++k) {
        double pw = 1.0;
        for (int e = 0; e < j - k; ++e) pw *= AB;
        out += binom * pw * G[i + k];
        binom = binom * (double)(j - k) / (double)(k + 1);
      }
      N[i][j] = out;
    }
}

__device__ inline void cxyz(int l, int i, int& x, int& y, int& z) {
  int xx = l, rem = i;
  while (rem > l - xx) {
    rem -= l - xx + 1;
    --xx;
  }
  x = xx;
  y = l - xx - rem;
  z = l - x - y;
}

__global__ void k_int1e_grad(const DevBasis b, const int* pstart, const int* pnum, const double* pexp,
                             const double* pcoef, int natm, const double* axyz, const double* aZ,
                             const int* shell_atom, const double* Dh, const double* Ws, double* grad) {
  const int sp = blockIdx.x * blockDim.x + threadIdx.x;
  if (sp >= b.nsp) return;
  const int A = b.sp_A[sp], B = b.sp_B[sp];
  const int la = b.shell_l[A], lb = b.shell_l[B];
  const int nA = ncart_c(la), nB = ncart_c(lb);
  const double* cA = b.shell_xyz + 3 * A;
  const double* cB = b.shell_xyz + 3 * B;
  const double ABv[3] = {cA[0] - cB[0], cA[1] - cB[1], cA[2] - cB[2]};
  const double AB2 = ABv[0] * ABv[0] + ABv[1] * ABv[1] + ABv[2] * ABv[2];
  const long long N = b.nao;
  const int a0 = b.shell_ao[A], b0 = b.shell_ao[B];
  const double fold = (A != B) ? 2.0 : 1.0;
  const int nr = (la + lb + 1) / 2 + 1;

  double gA[3] = {0, 0, 0}, gB[3] = {0, 0, 0};
  for (int ia_p = 0; ia_p < pnum[A]; ++ia_p)
    for (int ib_p = 0; ib_p < pnum[B]; ++ib_p) {
      const double al = pexp[pstart[A] + ia_p], be = pexp[pstart[B] + ib_p];
      const double c = pcoef[pstart[A] + ia_p] * pcoef[pstart[B] + ib_p];
      if (c == 0.0) continue;
      const double a2 = 2.0 * al, b2 = 2.0 * be;
      double S3[3][RI][RJ];
      for (int ax = 0; ax < 3; ++ax) ovl1d(la + 1, lb + 3, al, be, cA[ax], cB[ax], S3[ax]);
      for (int i = 0; i < nA; ++i) {
        int e[3];
        cxyz(la, i, e[0], e[1], e[2]);
        for (int j = 0; j < nB; ++j) {
          int f[3];
          cxyz(lb, j, f[0], f[1], f[2]);
          const double dh = Dh ? Dh[(a0 + i) * N + b0 + j] : 0.0;
          const double ws = Ws ? Ws[(a0 + i) * N + b0 + j] : 0.0;
          if (dh == 0.0 && ws == 0.0) continue;
          double s1[3], k1[3];
          for (int ax = 0; ax < 3; ++ax) {
            s1[ax] = S3[ax][e[ax]][f[ax]];
            k1[ax] = kin1(S3[ax], e[ax], f[ax], be);
          }
          for (int ax = 0; ax < 3; ++ax) {
            const int o1 = (ax + 1) % 3, o2 = (ax + 2) % 3;
            const double so = s1[o1] * s1[o2];                       // other two overlap factors
            const double ko = k1[o1] * s1[o2] + s1[o1] * k1[o2];     // kinetic on the other axes
            // A centre
            const double dsa = dA1(S3[ax], e[ax], f[ax], a2);
            const double dka = a2 * kin1(S3[ax], e[ax] + 1, f[ax], be) -
                               (e[ax] > 0 ? e[ax] * kin1(S3[ax], e[ax] - 1, f[ax], be) : 0.0);
            gA[ax] += c * fold * (dh * (dka * so + dsa * ko) - ws * dsa * so);
            // B centre
            const double dsb = dB1(S3[ax], e[ax], f[ax], b2);
            const double dkb = b2 * kin1(S3[ax], e[ax], f[ax] + 1, be) -
                               (f[ax] > 0 ? f[ax] * kin1(S3[ax], e[ax], f[ax] - 1, be) : 0.0);
            gB[ax] += c * fold * (dh * (dkb * so + dsb * ko) - ws * dsb * so);
          }
        }
      }
      if (!Dh || natm <= 0) continue;
      // nuclear attraction, one nucleus at a time
      const double p = al + be, ip = 1.0 / p;
      const double P[3] = {(al * cA[0] + be * cB[0]) * ip, (al * cA[1] + be * cB[1]) * ip,
                           (al * cA[2] + be * cB[2]) * ip};
      const double Kab = exp(-al * be * ip * AB2);
      for (int icn = 0; icn < natm; ++icn) {
        const double Z = aZ[icn];
        if (Z == 0.0) continue;
        const double PC[3] = {P[0] - axyz[3 * icn], P[1] - axyz[3 * icn + 1], P[2] - axyz[3 * icn + 2]};
        const double Targ = p * (PC[0] * PC[0] + PC[1] * PC[1] + PC[2] * PC[2]);
        const double pref = -Z * 2.0 * 3.14159265358979323846 * ip * Kab * c * fold;
        double vA[3] = {0, 0, 0}, vB[3] = {0, 0, 0};
        for (int u = 0; u < nr; ++u) {
          double xr, wr;
          rys_root_dev(nr, u, Targ, xr, wr);
          double N3[3][RI][RJ];
          for (int ax = 0; ax < 3; ++ax)
            nuc_tab(la + 1, lb + 1, P[ax] - cA[ax], PC[ax], ABv[ax], xr, 0.5 * ip, N3[ax]);
          for (int i = 0; i < nA; ++i) {
            int e[3];
            cxyz(la, i, e[0], e[1], e[2]);
            for (int j = 0; j < nB; ++j) {
              int f[3];
              cxyz(lb, j, f[0], f[1], f[2]);
              const double dh = Dh[(a0 + i) * N + b0 + j];
              if (dh == 0.0) continue;
              const double w = pref * wr * dh;
              for (int ax = 0; ax < 3; ++ax) {
                const int o1 = (ax + 1) % 3, o2 = (ax + 2) % 3;
                const double oth = N3[o1][e[o1]][f[o1]] * N3[o2][e[o2]][f[o2]];
                vA[ax] += w * dA1(N3[ax], e[ax], f[ax], a2) * oth;
                vB[ax] += w * dB1(N3[ax], e[ax], f[ax], b2) * oth;
              }
            }
          }
        }
        for (int ax = 0; ax < 3; ++ax) {
          gA[ax] += vA[ax];
          gB[ax] += vB[ax];
          atomicAdd(&grad[3 * icn + ax], -(vA[ax] + vB[ax]));
        }
      }
    }
  const int atA = shell_atom[A], atB = shell_atom[B];
  for (int ax = 0; ax < 3; ++ax) {
    atomicAdd(&grad[3 * atA + ax], gA[ax]);
    atomicAdd(&grad[3 * atB + ax], gB[ax]);
  }
}

}  // namespace

hipError_t launch_int1e_grad(const DevBasis& b, const int* pstart, const int* pnum, const double* pexp,
                             const double* pcoef, int natm, const double* axyz, const double* aZ,
                             const int* shell_atom, const double* Dh, const double* Ws, double* grad,
                             hipStream_t s) {
  hipLaunchKernelGGL(k_int1e_grad, dim3((b.nsp + 63) / 64), dim3(64), 0, s, b, pstart, pnum, pexp, pcoef, natm, axyz,
                     aZ, shell_atom, Dh, Ws, grad);
  return hipGetLastError();
}
