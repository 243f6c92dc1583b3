// Nuclear-derivative contraction of the four-centre ERIs: the backward pass of
// the J/K build.
//
// For symmetric adjoints W_J, W_K (upstream gradients of J and K) and the
// density D, the kernel accumulates
//   dL/dR  with  L = Tr(W_J J(D)) + Tr(W_K K(D)),
// holding D, W_J and W_K fixed, over the same canonical Schwarz-screened
// quartet list as the J/K build. For the RHF two-electron energy
// E_2e = 1/2 Tr(D J(D)) - 1/4 Tr(D K(D)) the adjoints are W_J = D/2,
// W_K = -D/4.
//
// Per canonical quartet (ab|cd) the adjoint tile mirrors the digestion of the
// forward build (the folded J/K updates, so sum_q sum_abcd bar (ab|cd) = L):
//   bar = fj (W_J[a,b] D[c,d] + [P!=Q] W_J[c,d] D[a,b])
//       + fk (W_K[a,c] D[b,d] + [C!=D] W_K[a,d] D[b,c]
//             + [A!=B] W_K[b,c] D[a,d] + [A!=B][C!=D] W_K[b,d] D[a,c]),
//   fj = (A!=B ? 2 : 1)(C!=D ? 2 : 1), fk = (P!=Q ? 2 : 1).
// The reference builds the same contraction over all ordered shell-pair pairs
// (asuka/mcscf/nac/_dense.py:195-250 for the adjoint tile, :1513-1839 for the
// loop); over canonical quartets with these factors it is the same sum.
//
// Derivatives of a primitive Cartesian Gaussian with respect to its centre
// (the reference's contraction kernel, asuka/cueri/_eri_rys_cpu.pyx:4316-4656):
//   d/dA_x (a b|c d) = 2 alpha (a+1_x b|c d) - a_x (a-1_x b|c d)
// and likewise for B (beta) and C (gamma); the D-centre term follows from
// translational invariance, dD = -(dA + dB + dC). The 2-D Rys tables are built
// to la+lb+1 and lc+ld+1 with floor((L+1)/2)+1 roots.
//
// One wavefront per quartet (the class pair's wave units, one ket at a time);
// LDS holds the roots, the VRR tables of all roots,
// the HRR-transferred 1-D tables of one root, and the adjoint tile of up to
// kChunk components with their table offsets. Lanes own components; the nine
// (centre A/B/C x xyz) sums are reduced across the wave at the end of the
// quartet and added to the atoms' gradient with FP64 atomics.
#include "jk_internal.h"
#include "rys_eval.h"

namespace {

constexpr int kWave = 64;
constexpr int kChunk = 1024;
constexpr int kGS = 8;  // VRR table stride (indices 0..7 = up to (f f) + 1)

// Binomial coefficients C(n, m), m <= n <= 4, as 3-bit fields of an immediate
// at bits 3(n(n+1)/2 + m): indexed per lane, a shift and a mask replace a
// memory load.
constexpr unsigned long long kBinom3 = 0x6684B64A249ull;
static_assert(((kBinom3 >> 36) & 7) == 6 && ((kBinom3 >> 24) & 7) == 3 && ((kBinom3 >> 42) & 7) == 1, "C(n, m)");
__device__ __forceinline__ double g_binom(int n, int m) {
  return (double)((kBinom3 >> (3 * (n * (n + 1) / 2 + m))) & 7ull);
}

// Component i of angular momentum l in the lx-descending, ly-descending order.
__device__ inline void cart_of(int l, int i, int& x, int& y) {
  int xx = l, rem = i;
  while (rem > l - xx) {
    rem -= l - xx + 1;
    --xx;
  }
  x = xx;
  y = l - xx - rem;
}

__device__ inline double wsum(double v) { return wave_total(v); }

struct GradArgs {
  const double* D;
  const double* WJ;  // may be null
  const double* WK;  // may be null
  const int* shell_atom;
  double* grad;  // (natm, 3)
};

// LDS layout (doubles unless noted), sized per class at launch:
//   rw[2*8] | G[8 roots][3][kGS*kGS] | I[3][NI] | bar[kChunk] | comp[kChunk][3] (int)
struct Lay {
  int nI, oG, oI, oBar, oComp, bytes;
};
__host__ __device__ inline Lay lay_of(int la, int lb, int lc, int ld) {
  Lay L;
  L.nI = (la + 2) * (lb + 2) * (lc + 2) * (ld + 1);
  L.oG = 16;
  L.oI = L.oG + 8 * 3 * kGS * kGS;
  L.oBar = L.oI + 3 * L.nI;
  L.oComp = L.oBar + kChunk;  // in doubles; comp ints start here
  L.bytes = (int)(sizeof(double) * L.oComp + sizeof(int) * 3 * kChunk);
  return L;
}

__global__ __launch_bounds__(kWave) void eri_deriv_kernel(DevBasis b, ClassWork w, GradArgs g) {
  extern __shared__ double lds[];
  const int lane = threadIdx.x;
  const int la = w.la, lb = w.lb, lc = w.lc, ld = w.ld;
  const Lay LY = lay_of(la, lb, lc, ld);
  double* s_rw = lds;
  double* s_G = lds + LY.oG;
  double* s_I = lds + LY.oI;
  double* s_bar = lds + LY.oBar;
  int* s_comp = reinterpret_cast<int*>(lds + LY.oComp);

  const int nA = ncart_c(la), nB = ncart_c(lb), nC = ncart_c(lc), nD = ncart_c(ld);
  const int nCD = nC * nD, ncomp = nA * nB * nCD;
  const int nr = (la + lb + lc + ld + 1) / 2 + 1;
  const int nmax = la + lb + 1, mmax = lc + ld + 1;
  const int sC = ld + 1, sB = (lc + 2) * sC, sA = (lb + 2) * sB;
  const long long N = b.nao;

  // item k: ket k % 64 of wave unit k / 64 (one quartet per wave at a time)
  for (long long k = blockIdx.x; k < w.nunits * 64; k += gridDim.x) {
    const WaveUnit un = w.units[k >> 6];
    if ((int)(k & 63) >= un.ncol) continue;  // uniform across the wave
    const int P = un.P, Qp = w.col_pair[un.col + (int)(k & 63)];
    const int A = b.sp_A[P], B = b.sp_B[P], C = b.sp_A[Qp], Dsh = b.sp_B[Qp];
    const double* cA = b.shell_xyz + 3 * A;
    const double* cB = b.shell_xyz + 3 * B;
    const double* cC = b.shell_xyz + 3 * C;
    const double* cD = b.shell_xyz + 3 * Dsh;
    const double AB[3] = {cA[0] - cB[0], cA[1] - cB[1], cA[2] - cB[2]};
    const double CD[3] = {cC[0] - cD[0], cC[1] - cD[1], cC[2] - cD[2]};
    const int aoA = b.shell_ao[A], aoB = b.shell_ao[B], aoC = b.shell_ao[C], aoD = b.shell_ao[Dsh];
    const bool ab_neq = A != B, cd_neq = C != Dsh, bk = P != Qp;
    const double fj = (ab_neq ? 2.0 : 1.0) * (cd_neq ? 2.0 : 1.0);
    const double fk = bk ? 2.0 : 1.0;
    const int ab0 = b.pp_start[P], nab = b.pp_count[P];
    const int cd0 = b.pp_start[Qp], ncd = b.pp_count[Qp];

    double gA[3] = {0, 0, 0}, gB[3] = {0, 0, 0}, gC[3] = {0, 0, 0};
    for (int c0 = 0; c0 < ncomp; c0 += kChunk) {
      const int nch = min(kChunk, ncomp - c0);
      __syncthreads();  // previous users of s_bar / s_comp are done
      for (int cc = lane; cc < nch; cc += kWave) {
        const int comp = c0 + cc;
        const int row = comp / nCD, col = comp - row * nCD;
        const int ia = row / nB, ib = row - ia * nB, ic = col / nD, id = col - ic * nD;
        const long long a = aoA + ia, bb = aoB + ib, c = aoC + ic, d = aoD + id;
        double v = 0.0;
        if (g.WJ) v += fj * (g.WJ[a * N + bb] * g.D[c * N + d] + (bk ? g.WJ[c * N + d] * g.D[a * N + bb] : 0.0));
        if (g.WK) {
          double t = g.WK[a * N + c] * g.D[bb * N + d];
          if (cd_neq) t += g.WK[a * N + d] * g.D[bb * N + c];
          if (ab_neq) t += g.WK[bb * N + c] * g.D[a * N + d];
          if (ab_neq && cd_neq) t += g.WK[bb * N + d] * g.D[a * N + c];
          v += fk * t;
        }
        s_bar[cc] = v;
        int ax, ay, bx, by, cx, cy, dx, dy;
        cart_of(la, ia, ax, ay);
        cart_of(lb, ib, bx, by);
        cart_of(lc, ic, cx, cy);
        cart_of(ld, id, dx, dy);
        const int az = la - ax - ay, bz = lb - bx - by, cz = lc - cx - cy, dz = ld - dx - dy;
        // per axis: table offset | a << 10 | b << 13 | c << 16
        s_comp[3 * cc + 0] = (ax * sA + bx * sB + cx * sC + dx) | (ax << 10) | (bx << 13) | (cx << 16);
        s_comp[3 * cc + 1] = (ay * sA + by * sB + cy * sC + dy) | (ay << 10) | (by << 13) | (cy << 16);
        s_comp[3 * cc + 2] = (az * sA + bz * sB + cz * sC + dz) | (az << 10) | (bz << 13) | (cz << 16);
      }

      for (int i = 0; i < nab; ++i) {
        const PrimPair X = b.pp[ab0 + i];
        for (int j = 0; j < ncd; ++j) {
          const PrimPair Y = b.pp[cd0 + j];
          const double p = X.eta, q = Y.eta, alpha2 = 2.0 * X.a_exp, beta2 = 2.0 * X.b_exp, gamma2 = 2.0 * Y.a_exp;
          const double inv_pq = 1.0 / (p + q);
          const double dPQ[3] = {X.Px - Y.Px, X.Py - Y.Py, X.Pz - Y.Pz};
          const double T = p * q * inv_pq * (dPQ[0] * dPQ[0] + dPQ[1] * dPQ[1] + dPQ[2] * dPQ[2]);
          const double base = 34.98683665524972497 * X.inv_eta * Y.inv_eta * sqrt(inv_pq) * X.cK * Y.cK;
          __syncthreads();  // previous primitive's tables are consumed
          if (lane < nr) {
            double xr, wr;
            rys_root_dev(nr, lane, T, xr, wr);
            s_rw[lane] = xr;
            s_rw[8 + lane] = wr;
          }
          __syncthreads();
          if (lane < 3 * nr) {  // VRR table of (root u, axis)
            const int u = lane / 3, ax = lane - 3 * u;
            const double x = s_rw[u];
            const double B0 = 0.5 * x * inv_pq;
            const double B1 = 0.5 * (1.0 - x) * X.inv_eta + B0;
            const double B1p = 0.5 * (1.0 - x) * Y.inv_eta + B0;
            const double PA = (ax == 0 ? X.Px : ax == 1 ? X.Py : X.Pz) - cA[ax];
            const double QC = (ax == 0 ? Y.Px : ax == 1 ? Y.Py : Y.Pz) - cC[ax];
            const double Cc = PA - q * inv_pq * x * dPQ[ax];
            const double Cp = QC + p * inv_pq * x * dPQ[ax];
            double* G = s_G + (u * 3 + ax) * kGS * kGS;
            G[0] = (ax == 2) ? s_rw[8 + u] * base : 1.0;
            if (nmax > 0) G[kGS] = Cc * G[0];
            for (int a = 2; a <= nmax; ++a) G[a * kGS] = B1 * (a - 1) * G[(a - 2) * kGS] + Cc * G[(a - 1) * kGS];
            for (int a = 0; a <= nmax; ++a) {
              if (mmax > 0) G[a * kGS + 1] = Cp * G[a * kGS] + (a > 0 ? a * B0 * G[(a - 1) * kGS] : 0.0);
              for (int c = 2; c <= mmax; ++c)
                G[a * kGS + c] = B1p * (c - 1) * G[a * kGS + c - 2] + Cp * G[a * kGS + c - 1] +
                                 (a > 0 ? a * B0 * G[(a - 1) * kGS + c - 1] : 0.0);
            }
          }
          for (int u = 0; u < nr; ++u) {
            __syncthreads();  // G ready / previous root's I consumed
            for (int e = lane; e < 3 * LY.nI; e += kWave) {
              const int ax = e / LY.nI, r = e - ax * LY.nI;
              const int ia = r / sA, r1 = r - ia * sA, ib = r1 / sB, r2 = r1 - ib * sB, ic = r2 / sC,
                        id = r2 - ic * sC;
              if (ia + ib > nmax) continue;  // never read
              const double* G = s_G + (u * 3 + ax) * kGS * kGS;
              double acc = 0.0, pab = 1.0;
              for (int kk = ib; kk >= 0; --kk) {  // sum_k C(ib,k) AB^(ib-k) ...
                double inner = 0.0, pcd = 1.0;
                for (int m = id; m >= 0; --m) {
                  inner += g_binom(id, m) * pcd * G[(ia + kk) * kGS + ic + m];
                  pcd *= CD[ax];
                }
                acc += g_binom(ib, kk) * pab * inner;
                pab *= AB[ax];
              }
              s_I[ax * LY.nI + r] = acc;
            }
            __syncthreads();
            const double* Ix = s_I;
            const double* Iy = s_I + LY.nI;
            const double* Iz = s_I + 2 * LY.nI;
            for (int cc = lane; cc < nch; cc += kWave) {
              const double bar = s_bar[cc];
              if (bar == 0.0) continue;
              const int px = s_comp[3 * cc], py = s_comp[3 * cc + 1], pz = s_comp[3 * cc + 2];
              const int ox = px & 1023, oy = py & 1023, oz = pz & 1023;
              const double x0 = Ix[ox], y0 = Iy[oy], z0 = Iz[oz];
              // d/dA, d/dB, d/dC of the 1-D factor of one axis
#define ASUKA_D1(T, o, pk, sh, stride, e2) \
  ((e2) * T[(o) + (stride)] - (((pk) >> (sh)) & 7 ? (double)(((pk) >> (sh)) & 7) * T[(o) - (stride)] : 0.0))
              const double yz = y0 * z0, xz = x0 * z0, xy = x0 * y0;
              gA[0] += bar * ASUKA_D1(Ix, ox, px, 10, sA, alpha2) * yz;
              gA[1] += bar * ASUKA_D1(Iy, oy, py, 10, sA, alpha2) * xz;
              gA[2] += bar * ASUKA_D1(Iz, oz, pz, 10, sA, alpha2) * xy;
              gB[0] += bar * ASUKA_D1(Ix, ox, px, 13, sB, beta2) * yz;
              gB[1] += bar * ASUKA_D1(Iy, oy, py, 13, sB, beta2) * xz;
              gB[2] += bar * ASUKA_D1(Iz, oz, pz, 13, sB, beta2) * xy;
              gC[0] += bar * ASUKA_D1(Ix, ox, px, 16, sC, gamma2) * yz;
              gC[1] += bar * ASUKA_D1(Iy, oy, py, 16, sC, gamma2) * xz;
              gC[2] += bar * ASUKA_D1(Iz, oz, pz, 16, sC, gamma2) * xy;
#undef ASUKA_D1
            }
          }
        }
      }
    }
    // reduce the nine sums over the wave; D from translational invariance
    double tot[9];
#pragma unroll
    for (int t = 0; t < 3; ++t) {
      tot[t] = wsum(gA[t]);
      tot[3 + t] = wsum(gB[t]);
      tot[6 + t] = wsum(gC[t]);
    }
    if (lane == 0) {
      const int atA = g.shell_atom[A], atB = g.shell_atom[B], atC = g.shell_atom[C], atD = g.shell_atom[Dsh];
#pragma unroll
      for (int t = 0; t < 3; ++t) {
        atomicAdd(&g.grad[3 * atA + t], tot[t]);
        atomicAdd(&g.grad[3 * atB + t], tot[3 + t]);
        atomicAdd(&g.grad[3 * atC + t], tot[6 + t]);
        atomicAdd(&g.grad[3 * atD + t], -(tot[t] + tot[3 + t] + tot[6 + t]));
      }
    }
  }
}

}  // namespace

size_t eri_deriv_lds_bytes(int la, int lb, int lc, int ld) { return (size_t)lay_of(la, lb, lc, ld).bytes; }

hipError_t launch_eri_deriv(const DevBasis& b, const ClassWork& w, const double* D, const double* WJ,
                            const double* WK, const int* shell_atom, double* grad, hipStream_t s) {
  if (w.nunits <= 0) return hipSuccess;
  GradArgs g{D, WJ, WK, shell_atom, grad};
  const size_t lds = eri_deriv_lds_bytes(w.la, w.lb, w.lc, w.ld);
  long long grid = w.nunits * 64 < 32768 ? w.nunits * 64 : 32768;
  hipLaunchKernelGGL(eri_deriv_kernel, dim3((unsigned)grid), dim3(kWave), lds, s, b, w, g);
  return hipGetLastError();
}
