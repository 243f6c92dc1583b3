// Wave-cooperative J/K kernels for the classes too large for the register
// kernels (d/f-heavy cc-pVTZ classes): one wavefront per shell quartet, class
// angular momenta as template parameters so every table size and index map is
// a compile-time constant.
//
// Work: the quartets of the class pair's wave units (a bra pair and up to 64
// kets of its row), one quartet per wave at a time.
// Per primitive quartet:
//   lanes u < NR          : Rys root/weight u from the interpolation table
//   lanes (u, axis) < 3NR : 2-D VRR table G[u][axis][i][k], built column by
//                           column in registers and stored to LDS (weight and
//                           prefactor folded into the z axis)
//   all lanes             : 1-D HRR tables I[u][axis][ia][ib][ic][id] via the
//                           binomial transfer from G
//   all lanes             : each lane adds sum_u Ix Iy Iz of its own Cartesian
//                           components into the quartet's tile in LDS
// After the primitive loop the tile is digested without atomics inside the
// wave: lanes own the elements of the six J/K blocks
// (cueri_cuda_kernels_dense_scatter.cu:246-352 semantics, symmetric updates
// folded) and contract the tile with the staged density blocks, then add each
// block element to J/K (or F) once.
#include "jk_internal.h"
#include "rys_eval.h"

namespace {

constexpr int kWave = 64;
constexpr int kPackMax = 16;  // registers per lane for packed component offsets

constexpr int ncc(int l) { return (l + 1) * (l + 2) / 2; }
// Cartesian exponents of l = 0..3 (offset l(l+1)(l+2)/6) and the binomial
// coefficients C(n, m), n, m <= 3, as 2-bit fields of immediates: exponent i
// at bits 2i, C(n, m) at bits 2(4n + m). Indexed per lane, a shift and a mask
// replace a memory load.
constexpr unsigned long long pack2(const signed char* v, int n) {
  unsigned long long r = 0;
  for (int i = 0; i < n; ++i) r |= (unsigned long long)v[i] << (2 * i);
  return r;
}
constexpr signed char k_lx[20] = {0, 1, 0, 0, 2, 1, 1, 0, 0, 0, 3, 2, 2, 1, 1, 1, 0, 0, 0, 0};
constexpr signed char k_ly[20] = {0, 0, 1, 0, 0, 1, 0, 2, 1, 0, 0, 1, 0, 2, 1, 0, 3, 2, 1, 0};
constexpr signed char k_lz[20] = {0, 0, 0, 1, 0, 0, 1, 0, 1, 2, 0, 0, 1, 0, 1, 2, 0, 1, 2, 3};
constexpr signed char k_binom[16] = {1, 0, 0, 0, 1, 1, 0, 0, 1, 2, 1, 0, 1, 3, 3, 1};
constexpr unsigned long long kLX = pack2(k_lx, 20), kLY = pack2(k_ly, 20), kLZ = pack2(k_lz, 20);
constexpr unsigned long long kBinom = pack2(k_binom, 16);
static_assert(((kLX >> 20) & 3) == 3 && ((kLY >> 32) & 3) == 3 && ((kLZ >> 38) & 3) == 3 &&
                  ((kBinom >> 26) & 3) == 3 && ((kBinom >> 18) & 3) == 2,
              "2-bit tables");
__device__ __forceinline__ int f2(unsigned long long t, int i) { return (int)((t >> (2 * i)) & 3ull); }

template <int LA, int LB, int LC, int LD>
struct CoopClass {
  static constexpr int NA = ncc(LA), NB = ncc(LB), NC = ncc(LC), ND = ncc(LD);
  static constexpr int NAB = NA * NB, NCD = NC * ND, NCOMP = NAB * NCD;
  static constexpr int NR = (LA + LB + LC + LD) / 2 + 1;
  static constexpr int NMAX = LA + LB, MMAX = LC + LD, GS = MMAX + 1, G1 = (NMAX + 1) * GS;
  // 1-D table strides: (ia, ib, ic, id)
  static constexpr int SIC = LD + 1, SIB = (LC + 1) * SIC, SIA = (LB + 1) * SIB, N1 = (LA + 1) * SIA;
  // LDS layout (doubles)
  static constexpr int O_RW = 0;                        // roots [NR], weights * prefactor [NR]
  static constexpr int O_G = 16;                        // [NR][3][G1]
  static constexpr int O_I = O_G + NR * 3 * G1;         // [NR][3][N1]
  static constexpr int O_T = O_I + NR * 3 * N1;         // tile [NCOMP]
  static constexpr int O_D = O_T + NCOMP;               // density blocks
  static constexpr int NDB = NCD + NAB + NB * ND + NB * NC + NA * ND + NA * NC;
  static constexpr int TOTAL = O_D + NDB;
  // outputs of the digestion: J_ab, J_cd, K_ac, K_ad, K_bc, K_bd
  static constexpr int NOUT = NAB + NCD + NA * NC + NA * ND + NB * NC + NB * ND;
};

template <int LA, int LB, int LC, int LD>
// At most 128 VGPRs (4 waves per SIMD); the few spilled values are outside the
// primitive loop.
__global__ __launch_bounds__(kWave) __attribute__((amdgpu_waves_per_eu(4))) void eri_coop_kernel(DevBasis b,
                                                                                                 ClassWork w,
                                                                                                 EriOut o) {
  using K = CoopClass<LA, LB, LC, LD>;
  constexpr int NA = K::NA, NB = K::NB, NC = K::NC, ND = K::ND, NAB = K::NAB, NCD = K::NCD;
  constexpr int NR = K::NR, NMAX = K::NMAX, MMAX = K::MMAX, GS = K::GS, G1 = K::G1;
  constexpr int SIA = K::SIA, SIB = K::SIB, SIC = K::SIC, N1 = K::N1;
  extern __shared__ double lds[];
  double* rw = lds + K::O_RW;
  double* G = lds + K::O_G;
  double* I = lds + K::O_I;
  double* tile = lds + K::O_T;
  double* Dcd = lds + K::O_D;
  double* Dab = Dcd + NCD;
  double* Dbd = Dab + NAB;
  double* Dbc = Dbd + NB * ND;
  double* Dad = Dbc + NB * NC;
  double* Dac = Dad + NA * ND;
  const int lane = threadIdx.x;
  constexpr int oA = LA * (LA + 1) * (LA + 2) / 6, oB = LB * (LB + 1) * (LB + 2) / 6;
  constexpr int oC = LC * (LC + 1) * (LC + 2) / 6, oD = LD * (LD + 1) * (LD + 2) / 6;

  // The lane's components and their 1-D table offsets (ix | iy << 10 | iz << 20)
  // are the same for every quartet: decoded once, kept in registers for the
  // classes with at most kPackMax components per lane.
  constexpr int NPL = (K::NCOMP + kWave - 1) / kWave;
  constexpr bool kPack = NPL <= kPackMax;
  static_assert(N1 <= 1024, "1-D table offsets are packed in 10 bits");
  int pk[kPack ? NPL : 1];
#pragma unroll
  for (int j = 0; j < (kPack ? NPL : 1); ++j) {
    const int comp = lane + kWave * j;
    int v = 0;
    if (kPack && comp < K::NCOMP) {
      const int iab = comp / NCD, icd = comp - iab * NCD;
      const int ia = iab / NB, ib = iab - ia * NB, ic = icd / ND, id = icd - ic * ND;
      const int ix = f2(kLX, oA + ia) * SIA + f2(kLX, oB + ib) * SIB + f2(kLX, oC + ic) * SIC + f2(kLX, oD + id);
      const int iy = f2(kLY, oA + ia) * SIA + f2(kLY, oB + ib) * SIB + f2(kLY, oC + ic) * SIC + f2(kLY, oD + id);
      const int iz = f2(kLZ, oA + ia) * SIA + f2(kLZ, oB + ib) * SIB + f2(kLZ, oC + ic) * SIC + f2(kLZ, oD + id);
      v = ix | (iy << 10) | (iz << 20);
    }
    pk[j] = v;
  }

  // item k: ket k % 64 of wave unit k / 64 (one quartet per wave at a time)
  for (long long k = blockIdx.x; k < w.nunits * 64; k += gridDim.x) {
    const WaveUnit un = w.units[k >> 6];
    const int jk = (int)(k & 63);
    if (jk >= un.ncol) continue;  // uniform: one quartet per wave
    const int P = un.P;
    const int A = b.sp_A[P], B = b.sp_B[P];
    const double* cA = b.shell_xyz + 3 * A;
    const double* cB = b.shell_xyz + 3 * B;
    const double AB[3] = {cA[0] - cB[0], cA[1] - cB[1], cA[2] - cB[2]};
    const int aoA = b.shell_ao[A], aoB = b.shell_ao[B];
    const int ab0 = b.pp_start[P], nab = b.pp_count[P];
    long long N, msh;  // the molecule's D/J/K block
    mol_layout(b, A, N, msh);
    const double* Dm = o.D + msh;
    {
      const int Qp = w.col_pair[un.col + jk];
      if (o.eps_density > 0.0 && (b.Q[P] * b.Q[Qp]) * fmax(b.Dsp[P], b.Dsp[Qp]) < o.eps_density) continue;
      const int C = b.sp_A[Qp], Dsh = b.sp_B[Qp];
      const double* cC = b.shell_xyz + 3 * C;
      const double* cD = b.shell_xyz + 3 * Dsh;
      const double CD[3] = {cC[0] - cD[0], cC[1] - cD[1], cC[2] - cD[2]};
      const int aoC = b.shell_ao[C], aoD = b.shell_ao[Dsh];
      const int cd0 = b.pp_start[Qp], ncd = b.pp_count[Qp];
      if (o.counters && lane == 0) {
        unsigned long long* cnt = o.counters + kCountStride * w.cp_index;
        atomicAdd(cnt + (((A != B) << 2) | ((C != Dsh) << 1) | (P != Qp)), 1ull);
        atomicAdd(cnt + 8, (unsigned long long)b.pp_count0[P] * (unsigned long long)b.pp_count0[Qp]);
      }
      __syncthreads();  // the previous quartet's digestion is done with the LDS
      for (int e = lane; e < K::NCOMP; e += kWave) tile[e] = 0.0;
      // density blocks of this quartet
      for (int e = lane; e < NCD; e += kWave) Dcd[e] = Dm[(aoC + e / ND) * N + aoD + e % ND];
      for (int e = lane; e < NAB; e += kWave) Dab[e] = Dm[(aoA + e / NB) * N + aoB + e % NB];
      for (int e = lane; e < NB * ND; e += kWave) Dbd[e] = Dm[(aoB + e / ND) * N + aoD + e % ND];
      for (int e = lane; e < NB * NC; e += kWave) Dbc[e] = Dm[(aoB + e / NC) * N + aoC + e % NC];
      for (int e = lane; e < NA * ND; e += kWave) Dad[e] = Dm[(aoA + e / ND) * N + aoD + e % ND];
      for (int e = lane; e < NA * NC; e += kWave) Dac[e] = Dm[(aoA + e / NC) * N + aoC + e % NC];

      for (int ipq = 0; ipq < nab * ncd; ++ipq) {
        const int iab_p = ipq / ncd, icd_p = ipq - iab_p * ncd;
        const PrimPair X = b.pp[ab0 + iab_p];
        const PrimPair Y = b.pp[cd0 + icd_p];
        const double p = X.eta, q = Y.eta;
        const double inv_pq = 1.0 / (p + q);
        const double dPQ[3] = {X.Px - Y.Px, X.Py - Y.Py, X.Pz - Y.Pz};
        const double T = p * q * inv_pq * (dPQ[0] * dPQ[0] + dPQ[1] * dPQ[1] + dPQ[2] * dPQ[2]);
        __syncthreads();  // the previous primitive's tables are consumed
        if (lane < NR) {
          double xr, wr;
          rys_root_dev(NR, lane, T, xr, wr);
          // 2 pi^(5/2) / (p q sqrt(p+q)) cK_ab cK_cd
          const double base = 34.98683665524972497 * X.inv_eta * Y.inv_eta * sqrt(inv_pq) * X.cK * Y.cK;
          rw[lane] = xr;
          rw[8 + lane] = wr * base;
        }
        __syncthreads();
        if (lane < 3 * NR) {  // VRR table of (root, axis), column by column in registers
          const int uu = lane / 3, ax = lane - 3 * uu;
          const double x = rw[uu];
          const double B0 = 0.5 * x * inv_pq;
          const double B1 = 0.5 * (1.0 - x) * X.inv_eta + B0;
          const double B1p = 0.5 * (1.0 - x) * Y.inv_eta + B0;
          const double Pa = ax == 0 ? X.Px : ax == 1 ? X.Py : X.Pz;
          const double Qa = ax == 0 ? Y.Px : ax == 1 ? Y.Py : Y.Pz;
          const double Cc = (Pa - cA[ax]) - q * inv_pq * x * dPQ[ax];
          const double Cp = (Qa - cC[ax]) + p * inv_pq * x * dPQ[ax];
          double* g = G + (uu * 3 + ax) * G1;
          double c0[NMAX + 1], c1[NMAX + 1], c2[NMAX + 1];  // columns k-2, k-1, k
          c1[0] = (ax == 2) ? rw[8 + uu] : 1.0;
          if (NMAX > 0) c1[1] = Cc * c1[0];
#pragma unroll
          for (int a = 2; a <= NMAX; ++a) c1[a] = B1 * (a - 1) * c1[a - 2] + Cc * c1[a - 1];
#pragma unroll
          for (int a = 0; a <= NMAX; ++a) {
            g[a * GS] = c1[a];
            c0[a] = 0.0;
          }
#pragma unroll
          for (int k = 1; k <= MMAX; ++k) {
#pragma unroll
            for (int a = 0; a <= NMAX; ++a) {
              double v = Cp * c1[a];
              if (k > 1) v += B1p * (k - 1) * c0[a];
              if (a > 0) v += a * B0 * c1[a - 1];
              c2[a] = v;
            }
#pragma unroll
            for (int a = 0; a <= NMAX; ++a) {
              g[a * GS + k] = c2[a];
              c0[a] = c1[a];
              c1[a] = c2[a];
            }
          }
        }
        __syncthreads();
        for (int e = lane; e < NR * 3 * N1; e += kWave) {  // HRR: G -> I
          const int uax = e / N1, r = e - uax * N1;
          const int ax = uax % 3;
          const int ia = r / SIA, r1 = r - ia * SIA;
          const int ib = r1 / SIB, r2 = r1 - ib * SIB;
          const int ic = r2 / SIC, id = r2 - ic * SIC;
          const double* g = G + uax * G1;
          double val = 0.0, pab = 1.0;
          for (int n = ib; n >= 0; --n) {  // sum_n C(ib,n) AB^(ib-n) sum_m C(id,m) CD^(id-m) G[ia+n][ic+m]
            double inner = 0.0, pcd = 1.0;
            for (int m = id; m >= 0; --m) {
              inner += (double)f2(kBinom, 4 * id + m) * pcd * g[(ia + n) * GS + ic + m];
              pcd *= CD[ax];
            }
            val += (double)f2(kBinom, 4 * ib + n) * pab * inner;
            pab *= AB[ax];
          }
          I[e] = val;
        }
        __syncthreads();
        if constexpr (kPack) {
#pragma unroll
          for (int j = 0; j < NPL; ++j) {  // this lane's components
            const int comp = lane + kWave * j;
            if (comp < K::NCOMP) {
              const int ix = pk[j] & 1023, iy = (pk[j] >> 10) & 1023, iz = pk[j] >> 20;
              double s = 0.0;
#pragma unroll
              for (int uu = 0; uu < NR; ++uu)
                s += I[(uu * 3) * N1 + ix] * I[(uu * 3 + 1) * N1 + iy] * I[(uu * 3 + 2) * N1 + iz];
              tile[comp] += s;
            }
          }
        } else
        for (int comp = lane; comp < K::NCOMP; comp += kWave) {  // this lane's components
          const int iab = comp / NCD, icd = comp - iab * NCD;
          const int ia = iab / NB, ib = iab - ia * NB, ic = icd / ND, id = icd - ic * ND;
          const int ix = f2(kLX, oA + ia) * SIA + f2(kLX, oB + ib) * SIB + f2(kLX, oC + ic) * SIC + f2(kLX, oD + id);
          const int iy = f2(kLY, oA + ia) * SIA + f2(kLY, oB + ib) * SIB + f2(kLY, oC + ic) * SIC + f2(kLY, oD + id);
          const int iz = f2(kLZ, oA + ia) * SIA + f2(kLZ, oB + ib) * SIB + f2(kLZ, oC + ic) * SIC + f2(kLZ, oD + id);
          double s = 0.0;
#pragma unroll
          for (int uu = 0; uu < NR; ++uu)
            s += I[(uu * 3) * N1 + ix] * I[(uu * 3 + 1) * N1 + iy] * I[(uu * 3 + 2) * N1 + iz];
          tile[comp] += s;
        }
      }
      __syncthreads();  // tile complete
      // ---- digestion: each lane owns block elements; no atomics inside the wave ----
      // (two-density builds digest the tile a second time with D2 into J2, K2)
      const bool ab_neq = A != B, cd_neq = C != Dsh, bk = P != Qp;
      const double fj = (ab_neq ? 2.0 : 1.0) * (cd_neq ? 2.0 : 1.0);
      const double fk = (bk ? 2.0 : 1.0) * (o.mode == MODE_FOCK ? -0.5 : 1.0);
      const int nden = o.D2 ? 2 : 1;
      for (int di = 0; di < nden; ++di) {
        double* Jo = shifted(di ? o.J2 : ((o.mode == MODE_FOCK) ? o.F : o.J), msh);
        double* Ko = shifted(di ? o.K2 : ((o.mode == MODE_FOCK) ? o.F : o.K), msh);
        if (di) {  // the density blocks of D2 replace those of D
          __syncthreads();
          const double* D2 = o.D2 + msh;
          for (int e = lane; e < NCD; e += kWave) Dcd[e] = D2[(aoC + e / ND) * N + aoD + e % ND];
          for (int e = lane; e < NAB; e += kWave) Dab[e] = D2[(aoA + e / NB) * N + aoB + e % NB];
          for (int e = lane; e < NB * ND; e += kWave) Dbd[e] = D2[(aoB + e / ND) * N + aoD + e % ND];
          for (int e = lane; e < NB * NC; e += kWave) Dbc[e] = D2[(aoB + e / NC) * N + aoC + e % NC];
          for (int e = lane; e < NA * ND; e += kWave) Dad[e] = D2[(aoA + e / ND) * N + aoD + e % ND];
          for (int e = lane; e < NA * NC; e += kWave) Dac[e] = D2[(aoA + e / NC) * N + aoC + e % NC];
          __syncthreads();
        }
        for (int e = lane; e < K::NOUT; e += kWave) {
          int r = e;
          if (r < NAB) {  // J[a,b] += fj sum_cd v D[c,d]
            if (!Jo) continue;
            double s = 0.0;
            for (int icd = 0; icd < NCD; ++icd) s += tile[r * NCD + icd] * Dcd[icd];
            acc_add(&Jo[(aoA + r / NB) * N + aoB + r % NB], fj * s, o.fixed);
            continue;
          }
          r -= NAB;
          if (r < NCD) {  // J[c,d] += fj sum_ab v D[a,b]
            if (!Jo || !bk) continue;
            double s = 0.0;
            for (int iab = 0; iab < NAB; ++iab) s += tile[iab * NCD + r] * Dab[iab];
            acc_add(&Jo[(aoC + r / ND) * N + aoD + r % ND], fj * s, o.fixed);
            continue;
          }
          r -= NCD;
          if (!Ko) continue;
          if (r < NA * NC) {  // K[a,c] += sum_{b,d} v D[b,d]
            const int ia = r / NC, ic = r - ia * NC;
            double s = 0.0;
            for (int ib = 0; ib < NB; ++ib)
              for (int id = 0; id < ND; ++id) s += tile[((ia * NB + ib) * NC + ic) * ND + id] * Dbd[ib * ND + id];
            acc_add(&Ko[(aoA + ia) * N + aoC + ic], fk * s, o.fixed);
            continue;
          }
          r -= NA * NC;
          if (r < NA * ND) {  // K[a,d] += sum_{b,c} v D[b,c]
            if (!cd_neq) continue;
            const int ia = r / ND, id = r - ia * ND;
            double s = 0.0;
            for (int ib = 0; ib < NB; ++ib)
              for (int ic = 0; ic < NC; ++ic) s += tile[((ia * NB + ib) * NC + ic) * ND + id] * Dbc[ib * NC + ic];
            acc_add(&Ko[(aoA + ia) * N + aoD + id], fk * s, o.fixed);
            continue;
          }
          r -= NA * ND;
          if (!ab_neq) continue;
          if (r < NB * NC) {  // K[b,c] += sum_{a,d} v D[a,d]
            const int ib = r / NC, ic = r - ib * NC;
            double s = 0.0;
            for (int ia = 0; ia < NA; ++ia)
              for (int id = 0; id < ND; ++id) s += tile[((ia * NB + ib) * NC + ic) * ND + id] * Dad[ia * ND + id];
            acc_add(&Ko[(aoB + ib) * N + aoC + ic], fk * s, o.fixed);
            continue;
          }
          r -= NB * NC;
          if (!cd_neq) continue;
          {  // K[b,d] += sum_{a,c} v D[a,c]
            const int ib = r / ND, id = r - ib * ND;
            double s = 0.0;
            for (int ia = 0; ia < NA; ++ia)
              for (int ic = 0; ic < NC; ++ic) s += tile[((ia * NB + ib) * NC + ic) * ND + id] * Dac[ia * NC + ic];
            acc_add(&Ko[(aoB + ib) * N + aoD + id], fk * s, o.fixed);
          }
        }
      }  // densities
    }
  }
}

template <int LA, int LB, int LC, int LD>
hipError_t launch_coop(const DevBasis& b, const ClassWork& w, const EriOut& o, hipStream_t s) {
  using K = CoopClass<LA, LB, LC, LD>;
  const size_t lds = sizeof(double) * K::TOTAL;
  long long grid = w.nunits * 64 < 32768 ? w.nunits * 64 : 32768;
  if (grid < 1) grid = 1;
  hipLaunchKernelGGL((eri_coop_kernel<LA, LB, LC, LD>), dim3((unsigned)grid), dim3(kWave), lds, s, b, w, o);
  return hipGetLastError();
}

using CoopLauncher = hipError_t (*)(const DevBasis&, const ClassWork&, const EriOut&, hipStream_t);

// Classes served here (LDS per wave <= about 48 KB); larger ones use the
// runtime-generic kernel.
#define ASUKA_COOP_LIST(X) \
  X(2, 2, 2, 1) X(2, 2, 2, 2) X(3, 1, 2, 2) X(3, 1, 3, 1) X(3, 2, 1, 1) X(3, 2, 2, 1) X(3, 2, 2, 2) X(3, 2, 3, 0) \
  X(3, 2, 3, 1) X(3, 3, 1, 1) X(3, 3, 2, 0) X(3, 3, 2, 1) X(3, 3, 3, 0) X(3, 3, 3, 1)

CoopLauncher find_coop(int la, int lb, int lc, int ld) {
#define ASUKA_COOP_CASE(a, b_, c, d) \
  if (la == a && lb == b_ && lc == c && ld == d) return &launch_coop<a, b_, c, d>;
  ASUKA_COOP_LIST(ASUKA_COOP_CASE)
#undef ASUKA_COOP_CASE
  return nullptr;
}

}  // namespace

hipError_t launch_eri_coop(const DevBasis& b, const ClassWork& w, const EriOut& o, hipStream_t s, bool* handled) {
  *handled = false;
  if (w.q_ab || !w.units || (o.mode != MODE_JK && o.mode != MODE_FOCK)) return hipSuccess;
  CoopLauncher f = find_coop(w.la, w.lb, w.lc, w.ld);
  if (!f) return hipSuccess;
  *handled = true;
  if (w.nunits <= 0) return hipSuccess;
  return f(b, w, o, s);
}
