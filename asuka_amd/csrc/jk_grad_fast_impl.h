// Lane-per-quartet nuclear-derivative contraction for the low angular-momentum
// classes (the bulk of the quartets), the backward pass of the J/K build.
//
// Same functional and adjoint tile as jk_grad.hip: per canonical quartet
//   bar = fj (W_J[a,b] D[c,d] + [P!=Q] W_J[c,d] D[a,b])
//       + fk (W_K[a,c] D[b,d] + [C!=D] W_K[a,d] D[b,c]
//             + [A!=B] W_K[b,c] D[a,d] + [A!=B][C!=D] W_K[b,d] D[a,c]),
// contracted with the centre derivatives of (ab|cd) per primitive quartet
// (asuka/cueri/_eri_rys_cpu.pyx:4316-4656):
//   d/dA_x (ab|cd) = 2 alpha (a+1_x b|cd) - a_x (a-1_x b|cd), likewise B, C;
//   d/dD = -(d/dA + d/dB + d/dC) (translational invariance).
//
// Work mapping as in the forward register kernels: workgroups pull row tasks
// (wave units of one bra pair) from the class pair's queue; a wave takes one
// unit, its lanes one ket each, so the bra pair (atoms of A and B, bra
// primitive pairs) is uniform across the wave. Per lane: the adjoint tile in
// registers; per primitive quartet and root the 2-D tables G_x, G_y, G_z up to
// (la+lb+1, lc+ld+1) in registers, and for every component the binomial
// transfer to the 1-D factors it needs (all indices compile-time). The A and B
// sums are reduced across the wave; all four centres' sums go to a per-
// workgroup LDS gradient (natm x 3), added to the output once per workgroup.
#pragma once
#include "jk_internal.h"
#include "rys_eval.h"

namespace asuka_gradf {

constexpr int kWaves = 4;
constexpr int kThreads = 64 * kWaves;
constexpr int kLdsGradMaxAtoms = 512;  // LDS gradient accumulator up to this many atoms

constexpr int ncg(int l) { return (l + 1) * (l + 2) / 2; }
constexpr int cgx(int l, int i) {
  int k = 0;
  for (int x = l; x >= 0; --x)
    for (int y = l - x; y >= 0; --y) {
      if (k == i) return x;
      ++k;
    }
  return 0;
}
constexpr int cgy(int l, int i) {
  int k = 0;
  for (int x = l; x >= 0; --x)
    for (int y = l - x; y >= 0; --y) {
      if (k == i) return y;
      ++k;
    }
  return 0;
}
constexpr int cge(int l, int i, int ax) { return ax == 0 ? cgx(l, i) : ax == 1 ? cgy(l, i) : l - cgx(l, i) - cgy(l, i); }
constexpr double binom_g(int n, int k) {
  double r = 1.0;
  for (int i = 1; i <= k; ++i) r = r * (n - k + i) / i;
  return r;
}

__device__ __forceinline__ double wave_sum_g(double v) { return wave_total(v); }
constexpr double kSqrt2Pi52g = 5.91496717279561287782;  // sqrt(2 pi^(5/2))
__device__ __forceinline__ int uni_g(int v) { return __builtin_amdgcn_readfirstlane(v); }

template <int LA, int LB, int LC, int LD>
struct DPlan {
  static constexpr int NA = ncg(LA), NB = ncg(LB), NC = ncg(LC), ND = ncg(LD);
  static constexpr int NCOMP = NA * NB * NC * ND;
  static constexpr int NR = (LA + LB + LC + LD + 1) / 2 + 1;
  static constexpr int NMAX = LA + LB + 1, MMAX = LC + LD + 1;
};

// 1-D factor of (i j | k l) on one axis from the VRR table G[n][m] (binomial
// transfer of the powers of j onto i and of l onto k); i..l compile-time.
template <int NMAX, int MMAX, int I, int J, int K, int L>
__device__ __forceinline__ double shift1(const double (&G)[NMAX + 1][MMAX + 1], const double* pab, const double* pcd) {
  double out = 0.0;
#pragma unroll
  for (int n = 0; n <= J; ++n) {
    double inner = 0.0;
#pragma unroll
    for (int m = 0; m <= L; ++m) inner += binom_g(L, m) * pcd[L - m] * G[I + n][K + m];
    out += binom_g(J, n) * pab[J - n] * inner;
  }
  return out;
}

template <int LA, int LB, int LC, int LD, int CI>
struct CompIdx {
  static constexpr int NB = ncg(LB), NC = ncg(LC), ND = ncg(LD);
  static constexpr int ia = CI / (NB * NC * ND), ib = (CI / (NC * ND)) % NB, ic = (CI / ND) % NC, id = CI % ND;
  static constexpr int a(int ax) { return cge(LA, ia, ax); }
  static constexpr int b(int ax) { return cge(LB, ib, ax); }
  static constexpr int c(int ax) { return cge(LC, ic, ax); }
  static constexpr int d(int ax) { return cge(LD, id, ax); }
};

template <int LA, int LB, int LC, int LD, int CI, int AX>
__device__ __forceinline__ void axis_vals(const double (&G)[LA + LB + 2][LC + LD + 2], const double* pab, const double* pcd,
                                          double a2, double b2, double c2, double& v0, double& vA, double& vB,
                                          double& vC) {
  using X = CompIdx<LA, LB, LC, LD, CI>;
  constexpr int NM = LA + LB + 1, MM = LC + LD + 1;
  constexpr int ea = X::a(AX), eb = X::b(AX), ec = X::c(AX), ed = X::d(AX);
  v0 = shift1<NM, MM, ea, eb, ec, ed>(G, pab, pcd);
  vA = a2 * shift1<NM, MM, ea + 1, eb, ec, ed>(G, pab, pcd);
  if constexpr (ea > 0) vA -= ea * shift1<NM, MM, (ea > 0 ? ea - 1 : 0), eb, ec, ed>(G, pab, pcd);
  vB = b2 * shift1<NM, MM, ea, eb + 1, ec, ed>(G, pab, pcd);
  if constexpr (eb > 0) vB -= eb * shift1<NM, MM, ea, (eb > 0 ? eb - 1 : 0), ec, ed>(G, pab, pcd);
  vC = c2 * shift1<NM, MM, ea, eb, ec + 1, ed>(G, pab, pcd);
  if constexpr (ec > 0) vC -= ec * shift1<NM, MM, ea, eb, (ec > 0 ? ec - 1 : 0), ed>(G, pab, pcd);
}

template <int LA, int LB, int LC, int LD, int CI, int C0, int C1>
__device__ __forceinline__ void comp_all(const double (&G)[3][LA + LB + 2][LC + LD + 2], const double (&pab)[3][LB + 2],
                                         const double (&pcd)[3][LD + 1], const double* bar, double a2, double b2,
                                         double c2, double (&gA)[3], double (&gB)[3], double (&gC)[3]) {
  if constexpr (CI < C1) {
    double v0[3], vA[3], vB[3], vC[3];
    axis_vals<LA, LB, LC, LD, CI, 0>(G[0], pab[0], pcd[0], a2, b2, c2, v0[0], vA[0], vB[0], vC[0]);
    axis_vals<LA, LB, LC, LD, CI, 1>(G[1], pab[1], pcd[1], a2, b2, c2, v0[1], vA[1], vB[1], vC[1]);
    axis_vals<LA, LB, LC, LD, CI, 2>(G[2], pab[2], pcd[2], a2, b2, c2, v0[2], vA[2], vB[2], vC[2]);
    const double w = bar[CI - C0];
    const double yz = w * v0[1] * v0[2], xz = w * v0[0] * v0[2], xy = w * v0[0] * v0[1];
    gA[0] = fma(vA[0], yz, gA[0]);
    gA[1] = fma(vA[1], xz, gA[1]);
    gA[2] = fma(vA[2], xy, gA[2]);
    gB[0] = fma(vB[0], yz, gB[0]);
    gB[1] = fma(vB[1], xz, gB[1]);
    gB[2] = fma(vB[2], xy, gB[2]);
    gC[0] = fma(vC[0], yz, gC[0]);
    gC[1] = fma(vC[1], xz, gC[1]);
    gC[2] = fma(vC[2], xy, gC[2]);
    comp_all<LA, LB, LC, LD, CI + 1, C0, C1>(G, pab, pcd, bar, a2, b2, c2, gA, gB, gC);
  }
}

struct GradFastArgs {
  const double* D;
  const double* WJ;  // may be null
  const double* WK;  // may be null
  const int* shell_atom;
  double* grad;  // (natm, 3)
  int natm;
};

// One wave unit, components [C0, C1) of the class (one chunk). pos: the
// lane's ket position (contiguous across the lanes; DevBasis::kpoff).
template <int LA, int LB, int LC, int LD, int C0, int C1>
__device__ __forceinline__ void deriv_unit(const DevBasis& b, const GradFastArgs& g, const double* tab, int P, int Qp,
                                           int pos, bool active, double* sgrad, bool lds_grad) {
  using DP = DPlan<LA, LB, LC, LD>;
  constexpr int NB = DP::NB, NC = DP::NC, ND = DP::ND;
  constexpr int NR = DP::NR, NM = DP::NMAX, MM = DP::MMAX;
  const int lane = threadIdx.x & 63;
  const long long N = b.nao;
  const int A = uni_g(b.sp_A[P]), B = uni_g(b.sp_B[P]), C = b.sp_A[Qp], Dsh = b.sp_B[Qp];
  const double* cA = b.shell_xyz + 3 * A;
  const double* cB = b.shell_xyz + 3 * B;
  const double* cC = b.shell_xyz + 3 * C;
  const double* cD = b.shell_xyz + 3 * Dsh;
  double pab[3][LB + 2], pcd[3][LD + 1];
#pragma unroll
  for (int ax = 0; ax < 3; ++ax) {
    const double ab = cA[ax] - cB[ax], cd = cC[ax] - cD[ax];
    pab[ax][0] = 1.0;
    pcd[ax][0] = 1.0;
#pragma unroll
    for (int e = 1; e <= LB + 1; ++e) pab[ax][e] = pab[ax][e - 1] * ab;
#pragma unroll
    for (int e = 1; e <= LD; ++e) pcd[ax][e] = pcd[ax][e - 1] * cd;
  }
  // adjoint tile of this lane's quartet (the chunk's components)
  const int aoA = b.shell_ao[A], aoB = b.shell_ao[B], aoC = b.shell_ao[C], aoD = b.shell_ao[Dsh];
  const bool ab_neq = A != B, cd_neq = C != Dsh, bk = P != Qp;
  const double fj = (ab_neq ? 2.0 : 1.0) * (cd_neq ? 2.0 : 1.0);
  const double fk = bk ? 2.0 : 1.0;
  double bar[C1 - C0];
#pragma unroll
  for (int ci = C0; ci < C1; ++ci) {
    const int ia = ci / (NB * NC * ND), ib = (ci / (NC * ND)) % NB, ic = (ci / ND) % NC, id = ci % ND;
    const long long a = aoA + ia, bb = aoB + ib, c = aoC + ic, d = aoD + id;
    double v = 0.0;
    if (active) {
      if (g.WJ) v += fj * (g.WJ[a * N + bb] * g.D[c * N + d] + (bk ? g.WJ[c * N + d] * g.D[a * N + bb] : 0.0));
      if (g.WK) {
        double s = g.WK[a * N + c] * g.D[bb * N + d];
        if (cd_neq) s += g.WK[a * N + d] * g.D[bb * N + c];
        if (ab_neq) s += g.WK[bb * N + c] * g.D[a * N + d];
        if (ab_neq && cd_neq) s += g.WK[bb * N + d] * g.D[a * N + c];
        v += fk * s;
      }
    }
    bar[ci - C0] = v;
  }
  double gA[3] = {0, 0, 0}, gB[3] = {0, 0, 0}, gC[3] = {0, 0, 0};
  const int ab0 = uni_g(b.pp_start[P]), nab = uni_g(b.pp_count[P]);
  const int ncd = active ? b.kpp[pos].y : 0;
  // the lane's ket primitive pairs from the transposed tables (slots
  // ks0 + 64 j, field f at kp64[f * kslots + slot]): one contiguous load
  // per field across the wave (no prefetch of the next pair: its registers
  // cost these kernels an occupancy step)
  const int ks0 = (int)b.kpoff[uni_g(pos >> 6)] * 64 + (pos & 63);
  const double* kt = b.kp64;
  const long long ns = b.kslots;
  auto ket = [&](int t, double* y) {
    y[0] = kt[t];           // eta
    y[1] = (kt + ns)[t];     // 1/eta
    y[2] = (kt + 2 * ns)[t]; // ck
    y[3] = (kt + 3 * ns)[t]; // P
    y[4] = (kt + 4 * ns)[t];
    y[5] = (kt + 5 * ns)[t];
    y[6] = (kt + 6 * ns)[t]; // P - C
    y[7] = (kt + 7 * ns)[t];
    y[8] = (kt + 8 * ns)[t];
    y[9] = (kt + 10 * ns)[t]; // exponent of the C primitive
  };
  for (int j = 0; j < ncd; ++j) {
    double y[10];
    ket(ks0 + 64 * j, y);
    const double q = y[0], iq = y[1], ckq = y[2], c2 = 2.0 * y[9];
    const double Yp[3] = {y[3], y[4], y[5]};
    const double QC[3] = {y[6], y[7], y[8]};
    for (int i = 0; i < nab; ++i) {
      const PrimPair X = b.pp[ab0 + i];  // wave-uniform
      const double p = X.eta, a2 = 2.0 * X.a_exp, b2 = 2.0 * X.b_exp;
      const double spq = p + q, inv_pq = rcp_nr(spq);
      const double dPQ[3] = {X.Px - Yp[0], X.Py - Yp[1], X.Pz - Yp[2]};
      const double T = p * q * inv_pq * (dPQ[0] * dPQ[0] + dPQ[1] * dPQ[1] + dPQ[2] * dPQ[2]);
      // 2 pi^(5/2) cK_ab cK_cd / (p q sqrt(p+q)); the ket's ck folds its share
      // (PrimPairF::ck = sqrt(2 pi^(5/2)) cK / eta)
      const double base = kSqrt2Pi52g * X.inv_eta * X.cK * ckq * rsq_nr(spq);
      const double PA[3] = {X.Px - cA[0], X.Py - cA[1], X.Pz - cA[2]};
      double rx[NR], rwt[NR];
      rys_roots_blk<NR, double, (NR <= 3 ? rys_c::lds_stride(NR, 8) : rys_c::interval_size(NR))>(T, tab, rx, rwt);
#pragma unroll
      for (int u = 0; u < NR; ++u) {
        const double x = rx[u];
        const double B0 = 0.5 * x * inv_pq;
        const double B1 = 0.5 * (1.0 - x) * X.inv_eta + B0;
        const double B1p = 0.5 * (1.0 - x) * iq + B0;
        const double xq = q * inv_pq * x, xp = p * inv_pq * x;
        double G[3][NM + 1][MM + 1];
#pragma unroll
        for (int ax = 0; ax < 3; ++ax) {
          const double Cc = PA[ax] - xq * dPQ[ax], Cp = QC[ax] + xp * dPQ[ax];
          G[ax][0][0] = (ax == 2) ? rwt[u] * base : 1.0;
          G[ax][1][0] = Cc * G[ax][0][0];
#pragma unroll
          for (int a = 2; a <= NM; ++a) G[ax][a][0] = B1 * (a - 1) * G[ax][a - 2][0] + Cc * G[ax][a - 1][0];
#pragma unroll
          for (int a = 0; a <= NM; ++a) {
            G[ax][a][1] = Cp * G[ax][a][0] + (a > 0 ? a * B0 * G[ax][a > 0 ? a - 1 : 0][0] : 0.0);
#pragma unroll
            for (int c = 2; c <= MM; ++c)
              G[ax][a][c] = B1p * (c - 1) * G[ax][a][c - 2] + Cp * G[ax][a][c - 1] +
                            (a > 0 ? a * B0 * G[ax][a > 0 ? a - 1 : 0][c - 1] : 0.0);
          }
        }
        comp_all<LA, LB, LC, LD, C0, C0, C1>(G, pab, pcd, bar, a2, b2, c2, gA, gB, gC);
      }
    }
  }
  // centres A and B are the same on every lane: reduce across the wave
  const int atA = g.shell_atom[A], atB = g.shell_atom[B];
#pragma unroll
  for (int ax = 0; ax < 3; ++ax) {
    const double sa = wave_sum_g(gA[ax]), sb = wave_sum_g(gB[ax]);
    if (lane == 0) {
      if (lds_grad) {
        atomicAdd(&sgrad[3 * atA + ax], sa);
        atomicAdd(&sgrad[3 * atB + ax], sb);
      } else {
        atomicAdd(&g.grad[3 * atA + ax], sa);
        atomicAdd(&g.grad[3 * atB + ax], sb);
      }
    }
  }
  if (active) {
    const int atC = g.shell_atom[C], atD = g.shell_atom[Dsh];
#pragma unroll
    for (int ax = 0; ax < 3; ++ax) {
      const double vd = -(gA[ax] + gB[ax] + gC[ax]);
      if (lds_grad) {
        atomicAdd(&sgrad[3 * atC + ax], gC[ax]);
        atomicAdd(&sgrad[3 * atD + ax], vd);
      } else {
        atomicAdd(&g.grad[3 * atC + ax], gC[ax]);
        atomicAdd(&g.grad[3 * atD + ax], vd);
      }
    }
  }
}

template <int LA, int LB, int LC, int LD, int NCH, int CH>
__device__ __forceinline__ void deriv_dispatch(int chunk, const DevBasis& b, const GradFastArgs& g, const double* tab,
                                               int P, int Qp, int pos, bool active, double* sgrad, bool lds_grad) {
  if constexpr (CH < NCH) {
    constexpr int NCOMP = DPlan<LA, LB, LC, LD>::NCOMP;
    if (chunk == CH) {
      deriv_unit<LA, LB, LC, LD, CH * NCOMP / NCH, (CH + 1) * NCOMP / NCH>(b, g, tab, P, Qp, pos, active, sgrad,
                                                                            lds_grad);
      return;
    }
    deriv_dispatch<LA, LB, LC, LD, NCH, CH + 1>(chunk, b, g, tab, P, Qp, pos, active, sgrad, lds_grad);
  }
}

// NCH: the class's components are split over NCH wave items per unit (each
// recomputes the primitive loop for its share), keeping tile + tables in
// registers.
template <int LA, int LB, int LC, int LD, int NCH>
__global__ __launch_bounds__(kThreads) void eri_deriv_fast_kernel(DevBasis b, ClassWork w, GradFastArgs g) {
  constexpr int NR = DPlan<LA, LB, LC, LD>::NR;
  constexpr bool kLds = NR <= 3;
  constexpr int TB = kLds ? rys_c::block_size(NR) : 1;
  constexpr int NFI = rys_c::interval_size(NR), S64 = rys_c::lds_stride(NR, 8);  // padded LDS intervals
  __shared__ double t64[kLds ? rys_c::nint[NR] * S64 : 1];
  __shared__ double sgrad[3 * kLdsGradMaxAtoms];
  __shared__ int s_task;
  const double* g64 = rys_dev::rys_coef + rys_c::coef_off[NR];
  if constexpr (kLds)
    for (int t = threadIdx.x; t < TB; t += kThreads) t64[(t / NFI) * S64 + t % NFI] = g64[t];
  const bool lds_grad = g.natm <= kLdsGradMaxAtoms;
  if (lds_grad)
    for (int t = threadIdx.x; t < 3 * g.natm; t += kThreads) sgrad[t] = 0.0;
  __syncthreads();
  const double* tab = kLds ? t64 : g64;
  const int lane = threadIdx.x & 63;
  const int wib = uni_g((int)(threadIdx.x >> 6));
  for (;;) {
    if (threadIdx.x == 0) s_task = (int)atomicAdd(w.queue, 1u);
    __syncthreads();
    const int t = s_task;
    if (t >= w.ntasks) break;
    const RowTask tk = w.tasks[t];
    for (int item = wib; item < tk.nu * NCH; item += kWaves) {  // chunk-major (instruction-cache locality)
      const int chunk = item / tk.nu, ui = item - chunk * tk.nu;
      const WaveUnit un = w.units[tk.u0 + ui];
      const int P = uni_g(un.P);
      const bool active = lane < un.ncol;
      const int Qp = w.col_pair[un.col + (active ? lane : 0)];
      // units start at 64-position blocks, so un.col + lane stays inside the
      // unit's block (padding positions for the inactive lanes)
      deriv_dispatch<LA, LB, LC, LD, NCH, 0>(chunk, b, g, tab, P, Qp, un.col + lane, active, sgrad, lds_grad);
    }
    __syncthreads();  // s_task may be overwritten
  }
  if (lds_grad) {
    for (int t = threadIdx.x; t < 3 * g.natm; t += kThreads) {
      const double v = sgrad[t];
      if (v != 0.0) atomicAdd(&g.grad[t], v);
    }
  }
}

template <int LA, int LB, int LC, int LD, int NCH>
hipError_t launch_deriv_fast(const DevBasis& b, const ClassWork& w, const GradFastArgs& g, hipStream_t s) {
  long long grid = w.ntasks < 2048 ? w.ntasks : 2048;
  if (grid < 1) grid = 1;
  hipLaunchKernelGGL((eri_deriv_fast_kernel<LA, LB, LC, LD, NCH>), dim3((unsigned)grid), dim3(kThreads), 0, s, b, w,
                     g);
  return hipGetLastError();
}

using DerivLauncher = hipError_t (*)(const DevBasis&, const ClassWork&, const GradFastArgs&, hipStream_t);

}  // namespace asuka_gradf

// Each jk_grad_fast_p*.hip translation unit instantiates a subset of classes:
// (la, lb, lc, ld, component chunks).
#define ASUKA_GRADF_FIND(FN, LIST)                                   \
  asuka_gradf::DerivLauncher FN(int la, int lb, int lc, int ld) {   \
    using namespace asuka_gradf;                                     \
    LIST(ASUKA_GRADF_CASE)                                           \
    return nullptr;                                                  \
  }
#define ASUKA_GRADF_CASE(a, b_, c, d, n) \
  if (la == a && lb == b_ && lc == c && ld == d) return &launch_deriv_fast<a, b_, c, d, n>;
