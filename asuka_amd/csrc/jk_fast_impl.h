// Class-specialised register kernels: one lane per shell quartet.
//
// For a class (la lb|lc ld) every lane loops over the primitive quartets of its
// shell quartet and, per Rys root, builds the 2-D tables
//   G_x[i][k], i <= la+lb, k <= lc+ld   (VRR, weight*prefactor folded into z)
// and accumulates the contracted intermediates
//   (e|f) = sum_prim sum_root G_x[ex][fx] G_y[ey][fy] G_z[ez][fz]
// for e in cart(la..la+lb), f in cart(lc..lc+ld), all in registers. The
// horizontal transfer to (ab|cd) runs once per shell quartet after the
// primitive loop (its coefficients depend only on A-B and C-D):
//   (e|cd) = sum_m prod_i C(d_i,m_i) CD_i^(d_i-m_i) (e|c+m)
//   (ab|cd) = sum_n prod_i C(b_i,n_i) AB_i^(b_i-n_i) (a+n|cd)
// The tile then feeds the J/K (or Fock) digestion of the reference's ordered
// contraction (cueri_cuda_kernels_dense_scatter.cu:246-352) with block sums
// per quartet: nAB + nCD J updates and up to 4 K blocks, instead of one atomic
// per tile element and term. Symmetric pair updates are folded into one (the
// outputs are symmetrised afterwards).
//
// All component indices are compile-time (constexpr tables + full unrolling),
// so the tables and accumulators live in VGPRs.
#pragma once
#include <type_traits>

#include "jk_internal.h"
#include "rys_eval.h"

namespace asuka_fast {

constexpr int nc_(int l) { return (l + 1) * (l + 2) / 2; }

// components of l in lx-descending, ly-descending order
constexpr int cx_(int l, int i) {
  int idx = 0;
  for (int x = l; x >= 0; --x)
    for (int y = l - x; y >= 0; --y) {
      if (idx == i) return x;
      ++idx;
    }
  return 0;
}
constexpr int cy_(int l, int i) {
  int idx = 0;
  for (int x = l; x >= 0; --x)
    for (int y = l - x; y >= 0; --y) {
      if (idx == i) return y;
      ++idx;
    }
  return 0;
}
constexpr int cidx_(int lx, int ly, int lz) {
  const int l = lx + ly + lz;
  return (l - lx) * (l - lx + 1) / 2 + lz;
}
constexpr int binom_(int n, int k) {
  int r = 1;
  for (int i = 1; i <= k; ++i) r = r * (n - k + i) / i;
  return r;
}

// Concatenated Cartesian range l in [lo, hi]: index -> exponents and back.
template <int LO, int HI>
struct Range {
  static constexpr int size() {
    int s = 0;
    for (int l = LO; l <= HI; ++l) s += nc_(l);
    return s;
  }
  static constexpr int off(int l) {
    int s = 0;
    for (int k = LO; k < l; ++k) s += nc_(k);
    return s;
  }
  struct Tab {
    int x[size()], y[size()], z[size()];
  };
  static constexpr Tab tab() {
    Tab t{};
    int i = 0;
    for (int l = LO; l <= HI; ++l)
      for (int c = 0; c < nc_(l); ++c) {
        t.x[i] = cx_(l, c);
        t.y[i] = cy_(l, c);
        t.z[i] = l - t.x[i] - t.y[i];
        ++i;
      }
    return t;
  }
  static constexpr int index(int x, int y, int z) { return off(x + y + z) + cidx_(x, y, z); }
};

template <typename CT>
__device__ __forceinline__ CT pw(const CT* p, int k) {
  return p[k];
}

// Plans split over A components whose wave items carry the J[c,d] sums
// across their A chunks (run_citem). Chosen by measurement at (H2O)8
// cc-pVTZ: for the other split plans the longer items (fewer per row task)
// and the extra live registers cost more than the atomics they save.
constexpr bool carry_jcd(int la, int lb, int lc, int ld) {
  const int k = ((la * 4 + lb) * 4 + lc) * 4 + ld;
  constexpr int on[] = {((2 * 4 + 1) * 4 + 1) * 4 + 1, ((2 * 4 + 1) * 4 + 2) * 4 + 1,
                        ((2 * 4 + 2) * 4 + 1) * 4 + 1, ((2 * 4 + 2) * 4 + 2) * 4 + 0,
                        ((3 * 4 + 0) * 4 + 2) * 4 + 1, ((3 * 4 + 0) * 4 + 2) * 4 + 2,
                        ((3 * 4 + 1) * 4 + 2) * 4 + 0, ((3 * 4 + 1) * 4 + 2) * 4 + 1,
                        ((3 * 4 + 2) * 4 + 2) * 4 + 0, ((3 * 4 + 3) * 4 + 1) * 4 + 0};
  for (int x : on)
    if (x == k) return true;
  return false;
}

template <int LA, int LB, int LC, int LD, int SA, int SC>
struct Plan {
  static constexpr int la = LA, lb = LB, lc = LC, ld = LD;
  static constexpr int NA = nc_(LA), NB = nc_(LB), NC = nc_(LC), ND = nc_(LD);
  static constexpr int NMAX = LA + LB, MMAX = LC + LD;
  static constexpr int NR = (LA + LB + LC + LD) / 2 + 1;
  static constexpr int SPLIT_A = SA, SPLIT_C = SC, NCH = SA * SC;
  static constexpr bool CARRY = SA > 1 && carry_jcd(LA, LB, LC, LD);
  using RE = Range<LA, LA + LB>;
  using RF = Range<LC, LC + LD>;
  static constexpr int NE = RE::size(), NF = RF::size();
};

// The chunk CH of a plan covers A-components [a0, a1) and C-components
// [c0, c1); it needs the intermediates (e|f) for e = a + n (|n| <= lb) and
// f = c + m (|m| <= ld) only.
template <class PL, int CH>
struct Chunk {
  static constexpr int ca = CH / PL::SPLIT_C, cc = CH % PL::SPLIT_C;
  static constexpr int a0 = ca * PL::NA / PL::SPLIT_A, a1 = (ca + 1) * PL::NA / PL::SPLIT_A;
  static constexpr int c0 = cc * PL::NC / PL::SPLIT_C, c1 = (cc + 1) * PL::NC / PL::SPLIT_C;
  struct Lists {
    int e[PL::NE];
    int emap[PL::NE];
    int ne;
    int f[PL::NF];
    int fmap[PL::NF];
    int nf;
  };
  static constexpr Lists lists() {
    Lists L{};
    bool ue[PL::NE] = {}, uf[PL::NF] = {};
    for (int a = a0; a < a1; ++a) {
      const int x = cx_(PL::la, a), y = cy_(PL::la, a), z = PL::la - x - y;
      for (int nx = 0; nx <= PL::lb; ++nx)
        for (int ny = 0; ny + nx <= PL::lb; ++ny)
          for (int nz = 0; nz + ny + nx <= PL::lb; ++nz) ue[PL::RE::index(x + nx, y + ny, z + nz)] = true;
    }
    for (int c = c0; c < c1; ++c) {
      const int x = cx_(PL::lc, c), y = cy_(PL::lc, c), z = PL::lc - x - y;
      for (int mx = 0; mx <= PL::ld; ++mx)
        for (int my = 0; my + mx <= PL::ld; ++my)
          for (int mz = 0; mz + my + mx <= PL::ld; ++mz) uf[PL::RF::index(x + mx, y + my, z + mz)] = true;
    }
    L.ne = 0;
    for (int i = 0; i < PL::NE; ++i) {
      L.emap[i] = -1;
      if (ue[i]) {
        L.e[L.ne] = i;
        L.emap[i] = L.ne++;
      }
    }
    L.nf = 0;
    for (int i = 0; i < PL::NF; ++i) {
      L.fmap[i] = -1;
      if (uf[i]) {
        L.f[L.nf] = i;
        L.fmap[i] = L.nf++;
      }
    }
    return L;
  }
  static constexpr Lists L = lists();
  static constexpr int NEL = L.ne, NFL = L.nf;
};

// Wave votes through the ballot intrinsic, whose result the compiler knows to
// be wave-uniform (branches on it stay scalar).
__device__ __forceinline__ bool wave_any(bool v) { return __builtin_amdgcn_ballot_w64(v) != 0ull; }
__device__ __forceinline__ bool wave_all(bool v) { return __builtin_amdgcn_ballot_w64(!v) == 0ull; }

// A value known to be the same on every lane, moved to a scalar register.
__device__ __forceinline__ int uni(int v) { return __builtin_amdgcn_readfirstlane(v); }

// Sum over the 64 lanes of the wavefront (all lanes must be converged).
__device__ __forceinline__ double wave_sum(double v) { return wave_total(v); }


// The ket primitive pair in slot t of the transposed tables (DevBasis::kpoff):
// the field bases are wave-uniform and the slot a 32-bit per-lane offset, so
// each field is one load with a scalar base and a vector offset.
template <typename PPT>
__device__ __forceinline__ PPT load_ket(const DevBasis& b, int t);
template <>
__device__ __forceinline__ PrimPairF load_ket<PrimPairF>(const DevBasis& b, int t) {
  const double* k = b.kp64;
  const long long n = b.kslots;
  PrimPairF y;
  y.eta = k[t];
  y.inv_eta = (k + n)[t];
  y.ck = (k + 2 * n)[t];
  y.Px = (k + 3 * n)[t];
  y.Py = (k + 4 * n)[t];
  y.Pz = (k + 5 * n)[t];
  y.PAx = (k + 6 * n)[t];
  y.PAy = (k + 7 * n)[t];
  y.PAz = (k + 8 * n)[t];
  y.q = (k + 9 * n)[t];
  return y;
}
template <>
__device__ __forceinline__ PrimPairF32 load_ket<PrimPairF32>(const DevBasis& b, int t) {
  const double* d = b.kp32d;
  const float* f = b.kp32f;
  const long long n = b.kslots;
  PrimPairF32 y;
  y.Px = d[t];
  y.Py = (d + n)[t];
  y.Pz = (d + 2 * n)[t];
  y.eta = f[t];
  y.inv_eta = (f + n)[t];
  y.ck = (f + 2 * n)[t];
  y.PAx = (f + 3 * n)[t];
  y.PAy = (f + 4 * n)[t];
  y.PAz = (f + 5 * n)[t];
  y.q = (f + 6 * n)[t];
  y.pad = 0.0f;
  return y;
}

// Arithmetic helpers for the compute type CT of the recursion (double, or
// float on the FP32 path).
__device__ __forceinline__ double fmac(double a, double b, double c) { return fma(a, b, c); }
__device__ __forceinline__ float fmac(float a, float b, float c) { return __builtin_fmaf(a, b, c); }

// Workgroups of up to 8 wavefronts share the root tables and the row
// accumulators; the host picks 4 or 8 per class pair (ClassWork::wpg): 8 when
// the accumulators are large (more waves per LDS byte), else 4 (less idle
// time at the end of short row tasks).
constexpr int kMaxWavesPerGroup = 8;
constexpr int kGroupThreads = 64 * kMaxWavesPerGroup;

// Workgroup accumulators of one row task (one bra pair P), in the dynamic
// LDS array s_acc: the J[a,b] block and the K rows of the bra's AOs, K[a, :]
// and K[b, :], over the columns of the ket class (AOs of angular momentum lc,
// then ld, in the compact per-l numbering). Offsets are in doubles; on ==
// false selects direct global atomics.
extern __shared__ double s_acc[];
struct RowAcc {
  bool on;
  int jab;  // [NA * NB]
  int kac;  // [NA * colsC]
  int kad;  // [NA * colsD] (same as kac when lc == ld)
  int kbc;  // [NB * colsC]
  int kbd;  // [NB * colsD] (same as kbc when lc == ld)
  int colsC, colsD;
  int stride;  // offset of the second density's accumulators (two-density builds)
};

// One chunk of one shell quartet per lane; the bra pair P is the same on every
// lane of the wave (bra-uniform wave units), so the bra primitive pairs are
// read with scalar loads in the inner loop while each lane streams its own
// ket primitive pairs in the outer loop. CT is the type of the primitive
// setup, roots, VRR, contracted intermediates and HRR; the digestion always
// accumulates in FP64. Geometric differences are formed in FP64 and rounded.
// `tab` is the NR-root block of the root table in CT (LDS or global).
// jcd: with the plan split over A components, the J[c,d] partial sums of the
// first density over this chunk's A components are added into jcd (the
// lane's registers, NCL*ND values) and the caller issues one atomic per
// element after the last A chunk; nullptr: atomics here.
// jb: the unit's block in the ket-position buffers (EriOut::jblk; uniform),
// or -1 for D[c,d] from D and J[c,d] atomics on J.

template <class PL, int CH, typename CT, bool FX>
__device__ __forceinline__ void run_chunk(const DevBasis& b, const EriOut& o, int P, int Qp, int pos, int nab,
                                          bool active, const CT* __restrict__ tab, const RowAcc& acc, double* jcd,
                                          long long jb) {
  using CK = Chunk<PL, CH>;
  constexpr int NEL = CK::NEL, NFL = CK::NFL, NR = PL::NR;
  constexpr int NMAX = PL::NMAX, MMAX = PL::MMAX;
  constexpr int LA = PL::la, LB = PL::lb, LC = PL::lc, LD = PL::ld;
  constexpr int NB = PL::NB, ND = PL::ND;
  constexpr int a0 = CK::a0, a1 = CK::a1, c0 = CK::c0, c1 = CK::c1;
  constexpr int NAL = a1 - a0, NCL = c1 - c0;
  constexpr int NCDL = NCL * ND;
  constexpr auto te = PL::RE::tab();
  constexpr auto tf = PL::RF::tab();
  constexpr auto L = CK::L;
  const CT one = CT(1), half = CT(0.5);

  // bra-side indices are wave-uniform; readfirstlane states it so that the
  // loads addressed by them are scalar
  // the ket pair's data by its position (pos: contiguous across the lanes)
  const int A = uni(b.sp_A[P]), B = uni(b.sp_B[P]);
  const double Ax = b.shell_xyz[3 * A], Ay = b.shell_xyz[3 * A + 1], Az = b.shell_xyz[3 * A + 2];
  const int ab0 = uni(b.pp_start[P]);
  const int2 kpp = b.kpp[pos];
  const int ncd = active ? kpp.y : 0;  // (kpp.x, the first pair in the pair table, is for host-side use)
  // the primitive pairs in the precision of CT (the FP32 path reads the float
  // copy: no conversions in the primitive loop but P - Q, formed in FP64)
  using PPT = typename std::conditional<sizeof(CT) == 4, PrimPairF32, PrimPairF>::type;
  const PPT* __restrict__ ppf;
  if constexpr (sizeof(CT) == 4)
    ppf = b.ppf32;
  else
    ppf = b.ppf;

  CT ef[NEL * NFL];
#pragma unroll
  for (int i = 0; i < NEL * NFL; ++i) ef[i] = CT(0);

  // Primitive-quartet screening: the primitive pairs of a shell pair are
  // sorted by their Schwarz bound q (largest first). nab (uniform, from the
  // wave unit) already leaves out the bra pairs whose q times the largest
  // ket-pair bound of the unit is below eps_prim; a lane leaves its ket loop
  // at the first ket pair whose q times the largest bra bound is below it.
  // Every skipped primitive quartet adds less than eps_prim to each integral.
  const CT kcut = (nab > 0 && o.eps_prim > 0.0) ? CT(o.eps_prim / (double)ppf[ab0].q) : CT(0);
  // the lane's ket primitive pairs, transposed by position (the lanes' j-th
  // pairs are contiguous): slots ks0 + 64 j
  const int ks0 = (int)b.kpoff[uni(pos >> 6)] * 64 + (pos & 63);
  PPT Yn;
  if (ncd > 0) Yn = load_ket<PPT>(b, ks0);
  for (int j = 0; j < ncd; ++j) {
    const PPT Y = Yn;
    if (Y.q < kcut) break;  // per lane: the rest of its ket pairs are smaller
    if (j + 1 < ncd) Yn = load_ket<PPT>(b, ks0 + 64 * (j + 1));  // next ket pair, in flight during this one
    const CT q = CT(Y.eta), iq = CT(Y.inv_eta), ckq = CT(Y.ck);
    const CT QCx = CT(Y.PAx), QCy = CT(Y.PAy), QCz = CT(Y.PAz);
    for (int i = 0; i < nab; ++i) {
      const PPT X = ppf[ab0 + i];  // wave-uniform
      const CT p = CT(X.eta), ip = CT(X.inv_eta);
      const CT PAx = CT(X.PAx), PAy = CT(X.PAy), PAz = CT(X.PAz);
      const CT spq = p + q;
      const CT inv_pq = rcp_nr(spq);
      const CT dx = CT(X.Px - Y.Px), dy = CT(X.Py - Y.Py), dz = CT(X.Pz - Y.Pz);
      const CT T = p * q * inv_pq * (dx * dx + dy * dy + dz * dz);
      // 2 pi^(5/2) cK_ab cK_cd / (p q sqrt(p+q))
      const CT base = CT(X.ck) * ckq * rsq_nr(spq);
      CT rx[NR], rw[NR];
      rys_roots_blk<NR, CT, (NR <= kLdsRootsMax ? rys_c::lds_stride_of<CT>(NR) : rys_c::interval_size_of<CT>(NR))>(
          T, tab, rx, rw);
      const CT qf = q * inv_pq, pf = p * inv_pq;
#pragma unroll
      for (int u = 0; u < NR; ++u) {
        const CT x = rx[u];
        const CT B0 = half * x * inv_pq;
        const CT B1 = half * (one - x) * ip + B0;
        const CT B1p = half * (one - x) * iq + B0;
        const CT xq = qf * x, xp = pf * x;
        CT G[3][NMAX + 1][MMAX + 1];
        const CT c00[3] = {PAx - xq * dx, PAy - xq * dy, PAz - xq * dz};
        const CT c0p[3] = {QCx + xp * dx, QCy + xp * dy, QCz + xp * dz};
#pragma unroll
        for (int ax = 0; ax < 3; ++ax) {
          const CT g0 = (ax == 2) ? rw[u] * base : one;
          const CT Cc = c00[ax], Cp = c0p[ax];
          G[ax][0][0] = g0;
          if (NMAX > 0) G[ax][1][0] = Cc * g0;
#pragma unroll
          for (int a = 2; a <= NMAX; ++a) G[ax][a][0] = B1 * CT(a - 1) * G[ax][a - 2][0] + Cc * G[ax][a - 1][0];
          if (MMAX > 0) {
#pragma unroll
            for (int a = 0; a <= NMAX; ++a) {
              G[ax][a][1] = Cp * G[ax][a][0] + (a > 0 ? CT(a) * B0 * G[ax][a > 0 ? a - 1 : 0][0] : CT(0));
#pragma unroll
              for (int c = 2; c <= MMAX; ++c)
                G[ax][a][c] = B1p * CT(c - 1) * G[ax][a][c - 2] + Cp * G[ax][a][c - 1] +
                              (a > 0 ? CT(a) * B0 * G[ax][a > 0 ? a - 1 : 0][c - 1] : CT(0));
            }
          }
        }
#pragma unroll
        for (int ie = 0; ie < NEL; ++ie) {
#pragma unroll
          for (int jf = 0; jf < NFL; ++jf) {
            const int eg = L.e[ie], fg = L.f[jf];
            const CT gxy = G[0][te.x[eg]][tf.x[fg]] * G[1][te.y[eg]][tf.y[fg]];
            ef[ie * NFL + jf] = fmac(gxy, G[2][te.z[eg]][tf.z[fg]], ef[ie * NFL + jf]);
          }
        }
      }
    }
  }

  // ---- horizontal transfer: ket (C-components of the chunk) then bra ----
  const double Bx = b.shell_xyz[3 * B], By = b.shell_xyz[3 * B + 1], Bz = b.shell_xyz[3 * B + 2];
  CT pab[3][LB + 1], pcd[3][LD + 1];
  {
    const CT ab[3] = {CT(Ax - Bx), CT(Ay - By), CT(Az - Bz)};
    const CT cd[3] = {CT(b.kcd[3 * pos]), CT(b.kcd[3 * pos + 1]), CT(b.kcd[3 * pos + 2])};
#pragma unroll
    for (int ax = 0; ax < 3; ++ax) {
      pab[ax][0] = one;
      pcd[ax][0] = one;
#pragma unroll
      for (int e = 1; e <= LB; ++e) pab[ax][e] = pab[ax][e - 1] * ab[ax];
#pragma unroll
      for (int e = 1; e <= LD; ++e) pcd[ax][e] = pcd[ax][e - 1] * cd[ax];
    }
  }
  CT ecd[NEL * NCDL];
#pragma unroll
  for (int ie = 0; ie < NEL; ++ie)
#pragma unroll
    for (int ic = 0; ic < NCL; ++ic)
#pragma unroll
      for (int id = 0; id < ND; ++id) {
        const int cxv = cx_(LC, c0 + ic), cyv = cy_(LC, c0 + ic), czv = LC - cxv - cyv;
        const int dxv = cx_(LD, id), dyv = cy_(LD, id), dzv = LD - dxv - dyv;
        CT s = CT(0);
#pragma unroll
        for (int mx = 0; mx <= dxv; ++mx)
#pragma unroll
          for (int my = 0; my <= dyv; ++my)
#pragma unroll
            for (int mz = 0; mz <= dzv; ++mz) {
              const CT coef = CT(binom_(dxv, mx) * binom_(dyv, my) * binom_(dzv, mz)) *
                                  pw(pcd[0], dxv - mx) * pw(pcd[1], dyv - my) * pw(pcd[2], dzv - mz);
              s = fmac(coef, ef[ie * NFL + L.fmap[PL::RF::index(cxv + mx, cyv + my, czv + mz)]], s);
            }
        ecd[ie * NCDL + ic * ND + id] = s;
      }
  CT tile[NAL * NB * NCDL];
#pragma unroll
  for (int ia = 0; ia < NAL; ++ia)
#pragma unroll
    for (int ib = 0; ib < NB; ++ib) {
      const int axv = cx_(LA, a0 + ia), ayv = cy_(LA, a0 + ia), azv = LA - axv - ayv;
      const int bxv = cx_(LB, ib), byv = cy_(LB, ib), bzv = LB - bxv - byv;
#pragma unroll
      for (int jc = 0; jc < NCDL; ++jc) {
        CT s = CT(0);
#pragma unroll
        for (int nx = 0; nx <= bxv; ++nx)
#pragma unroll
          for (int ny = 0; ny <= byv; ++ny)
#pragma unroll
            for (int nz = 0; nz <= bzv; ++nz) {
              const CT coef = CT(binom_(bxv, nx) * binom_(byv, ny) * binom_(bzv, nz)) *
                                  pw(pab[0], bxv - nx) * pw(pab[1], byv - ny) * pw(pab[2], bzv - nz);
              s = fmac(coef, ecd[L.emap[PL::RE::index(axv + nx, ayv + ny, azv + nz)] * NCDL + jc], s);
            }
        tile[(ia * NB + ib) * NCDL + jc] = s;
      }
    }

  // ---- digestion of the chunk's tile block (rows a0..a1 x all b, cols c0..c1 x all d) ----
  // Sums over the chunk's own components are complete; the rest are partial
  // and add up across chunks. J[a,b] and the K rows of a and b go to the
  // workgroup's LDS accumulators when present (flushed once per row task),
  // J[c,d] (distinct per lane) to global atomics. With a second density
  // (two-density builds) the same tile is digested again into (J2, K2).
  long long N, msh;  // the molecule's D/J/K block (uniform: the bra pair is)
  mol_layout(b, A, N, msh);
  const int2 kao = b.kao[pos];
  const int aoA = b.shell_ao[A] + a0, aoB = b.shell_ao[B], aoC = kao.x + c0, aoD = kao.y;
  const bool ab_neq = A != B, cd_neq = kao.x != kao.y, bk = P != Qp;
  const double fj = (ab_neq ? 2.0 : 1.0) * (cd_neq ? 2.0 : 1.0);
  const double fk = bk ? 2.0 : 1.0;
  const double ks = (o.mode == MODE_FOCK) ? -0.5 * fk : fk;
  const bool lds = acc.on;  // uniform
  // compact columns of the lane's C (first component of the chunk) and D shells
  const int2 klao = b.klao[pos];
  const int colC = lds ? klao.x + c0 : 0, colD = lds ? klao.y : 0;
  const int nden = o.D2 ? 2 : 1;  // uniform
  for (int di = 0; di < nden; ++di) {
    const double* Dm = (di ? o.D2 : o.D) + msh;
    double* Jo = shifted(di ? o.J2 : ((o.mode == MODE_FOCK) ? o.F : o.J), msh);
    double* Ko = shifted(di ? o.K2 : ((o.mode == MODE_FOCK) ? o.F : o.K), msh);
    const int so = di * acc.stride;  // this density's accumulators in s_acc
    if (Jo) {
      const int lane = threadIdx.x & 63;
      double dcd[NCDL];
      if (jb >= 0) {  // the chunk's components c0 * ND.. of the lane's ket position
        const double* Dp = (di ? o.Dpos2 : o.Dpos) + jb + 64LL * (c0 * ND) + lane;
#pragma unroll
        for (int c = 0; c < NCDL; ++c) dcd[c] = Dp[64 * c];
      } else {
#pragma unroll
        for (int ic = 0; ic < NCL; ++ic)
#pragma unroll
          for (int id = 0; id < ND; ++id) dcd[ic * ND + id] = Dm[(aoC + ic) * N + aoD + id];
      }
      // The bra pair, hence the J[a,b] addresses, is the same on every lane:
      // reduce across the wave, one update per element.
      const unsigned long long amask = __ballot(active);
      const int first = amask ? (__ffsll((unsigned long long)amask) - 1) : 0;
#pragma unroll
      for (int r = 0; r < NAL * NB; ++r) {
        double s = 0.0;
#pragma unroll
        for (int c = 0; c < NCDL; ++c) s = fma((double)tile[r * NCDL + c], dcd[c], s);
        s = wave_sum(active ? fj * s : 0.0);
        if (lane == first && active) {
          if (lds)
            acc_add(&s_acc[so + acc.jab + (a0 + r / NB) * NB + r % NB], s, FX);
          else
            acc_add(&Jo[(aoA + r / NB) * N + aoB + r % NB], s, FX);
        }
      }
      if (bk && active) {
        double dab[NAL * NB];
#pragma unroll
        for (int ia = 0; ia < NAL; ++ia)
#pragma unroll
          for (int ib = 0; ib < NB; ++ib) dab[ia * NB + ib] = Dm[(aoA + ia) * N + aoB + ib];
#pragma unroll
        for (int c = 0; c < NCDL; ++c) {
          double s = 0.0;
#pragma unroll
          for (int r = 0; r < NAL * NB; ++r) s = fma((double)tile[r * NCDL + c], dab[r], s);
          if (jcd && di == 0)
            jcd[c] = fma(fj, s, jcd[c]);
          else if (jb >= 0)
            acc_add((di ? o.Jpos2 : o.Jpos) + jb + 64LL * (c0 * ND + c) + lane, fj * s, FX);
          else
            acc_add(&Jo[(aoC + c / ND) * N + aoD + c % ND], fj * s, FX);
        }
      }
    }
    if (Ko && active) {
      {  // K[a,c] += sum_{b,d} v D[b,d]
        double dbd[NB * ND];
#pragma unroll
        for (int ib = 0; ib < NB; ++ib)
#pragma unroll
          for (int id = 0; id < ND; ++id) dbd[ib * ND + id] = Dm[(aoB + ib) * N + aoD + id];
#pragma unroll
        for (int ia = 0; ia < NAL; ++ia)
#pragma unroll
          for (int ic = 0; ic < NCL; ++ic) {
            double s = 0.0;
#pragma unroll
            for (int ib = 0; ib < NB; ++ib)
#pragma unroll
              for (int id = 0; id < ND; ++id) s = fma((double)tile[(ia * NB + ib) * NCDL + ic * ND + id], dbd[ib * ND + id], s);
            if (lds)
              acc_add(&s_acc[so + acc.kac + (a0 + ia) * acc.colsC + colC + ic], ks * s, FX);
            else
              acc_add(&Ko[(aoA + ia) * N + aoC + ic], ks * s, FX);
          }
      }
      if (cd_neq) {  // K[a,d] += sum_{b,c} v D[b,c]
        double dbc[NB * NCL];
#pragma unroll
        for (int ib = 0; ib < NB; ++ib)
#pragma unroll
          for (int ic = 0; ic < NCL; ++ic) dbc[ib * NCL + ic] = Dm[(aoB + ib) * N + aoC + ic];
#pragma unroll
        for (int ia = 0; ia < NAL; ++ia)
#pragma unroll
          for (int id = 0; id < ND; ++id) {
            double s = 0.0;
#pragma unroll
            for (int ib = 0; ib < NB; ++ib)
#pragma unroll
              for (int ic = 0; ic < NCL; ++ic) s = fma((double)tile[(ia * NB + ib) * NCDL + ic * ND + id], dbc[ib * NCL + ic], s);
            if (lds)
              acc_add(&s_acc[so + acc.kad + (a0 + ia) * acc.colsD + colD + id], ks * s, FX);
            else
              acc_add(&Ko[(aoA + ia) * N + aoD + id], ks * s, FX);
          }
      }
      if (ab_neq) {  // K[b,c] += sum_{a,d} v D[a,d]
        double dad[NAL * ND];
#pragma unroll
        for (int ia = 0; ia < NAL; ++ia)
#pragma unroll
          for (int id = 0; id < ND; ++id) dad[ia * ND + id] = Dm[(aoA + ia) * N + aoD + id];
#pragma unroll
        for (int ib = 0; ib < NB; ++ib)
#pragma unroll
          for (int ic = 0; ic < NCL; ++ic) {
            double s = 0.0;
#pragma unroll
            for (int ia = 0; ia < NAL; ++ia)
#pragma unroll
              for (int id = 0; id < ND; ++id) s = fma((double)tile[(ia * NB + ib) * NCDL + ic * ND + id], dad[ia * ND + id], s);
            if (lds)
              acc_add(&s_acc[so + acc.kbc + ib * acc.colsC + colC + ic], ks * s, FX);
            else
              acc_add(&Ko[(aoB + ib) * N + aoC + ic], ks * s, FX);
          }
        if (cd_neq) {  // K[b,d] += sum_{a,c} v D[a,c]
          double dac[NAL * NCL];
#pragma unroll
          for (int ia = 0; ia < NAL; ++ia)
#pragma unroll
            for (int ic = 0; ic < NCL; ++ic) dac[ia * NCL + ic] = Dm[(aoA + ia) * N + aoC + ic];
#pragma unroll
          for (int ib = 0; ib < NB; ++ib)
#pragma unroll
            for (int id = 0; id < ND; ++id) {
              double s = 0.0;
#pragma unroll
              for (int ia = 0; ia < NAL; ++ia)
#pragma unroll
                for (int ic = 0; ic < NCL; ++ic)
                  s = fma((double)tile[(ia * NB + ib) * NCDL + ic * ND + id], dac[ia * NCL + ic], s);
              if (lds)
                acc_add(&s_acc[so + acc.kbd + ib * acc.colsD + colD + id], ks * s, FX);
              else
                acc_add(&Ko[(aoB + ib) * N + aoD + id], ks * s, FX);
            }
        }
      }
    }
  }  // densities
}

template <class PL, int CH, typename CT, bool FX>
__device__ __forceinline__ void dispatch_chunk(int chunk, const DevBasis& b, const EriOut& o, int P, int Qp, int pos,
                                               int nab, bool active, const CT* __restrict__ tab, const RowAcc& acc,
                                               long long jb) {
  if constexpr (CH < PL::NCH) {
    if (chunk == CH) {
      run_chunk<PL, CH, CT, FX>(b, o, P, Qp, pos, nab, active, tab, acc, nullptr, jb);
      return;
    }
    dispatch_chunk<PL, CH + 1, CT, FX>(chunk, b, o, P, Qp, pos, nab, active, tab, acc, jb);
  }
}

// The A chunks CA.. of C chunk CC, in sequence, carrying J[c,d] in jcd.
template <class PL, int CC, int CA, typename CT, bool FX>
__device__ __forceinline__ void run_agroup(const DevBasis& b, const EriOut& o, int P, int Qp, int pos, int nab,
                                           bool active, const CT* __restrict__ tab, const RowAcc& acc, double* jcd,
                                           long long jb) {
  if constexpr (CA < PL::SPLIT_A) {
    run_chunk<PL, CA * PL::SPLIT_C + CC, CT, FX>(b, o, P, Qp, pos, nab, active, tab, acc, jcd, jb);
    run_agroup<PL, CC, CA + 1, CT, FX>(b, o, P, Qp, pos, nab, active, tab, acc, jcd, jb);
  }
}

// One wave item of a CARRY plan: C chunk CC of a unit, all of its A chunks,
// with the J[c,d] sums carried across the A chunks in registers, so J[c,d]
// receives one atomic per element and quartet instead of one per A chunk.
template <class PL, int CC, typename CT, bool FX>
__device__ __forceinline__ void run_citem(const DevBasis& b, const EriOut& o, int P, int Qp, int pos, int nab,
                                          bool active, const CT* __restrict__ tab, const RowAcc& acc, long long jb) {
  using CK = Chunk<PL, CC>;  // A chunk 0: the C range of every A chunk of CC
  constexpr int ND = PL::ND, NCDL = (CK::c1 - CK::c0) * ND;
  double jcd[NCDL];
#pragma unroll
  for (int c = 0; c < NCDL; ++c) jcd[c] = 0.0;
  run_agroup<PL, CC, 0, CT, FX>(b, o, P, Qp, pos, nab, active, tab, acc, jcd, jb);
  double* Jo = (o.mode == MODE_FOCK) ? o.F : o.J;
  if (Jo && active && P != Qp && jb >= 0) {
    const int lane = threadIdx.x & 63;
#pragma unroll
    for (int c = 0; c < NCDL; ++c) acc_add(o.Jpos + jb + 64LL * (CK::c0 * ND + c) + lane, jcd[c], FX);
  } else if (Jo && active && P != Qp) {
    const int A = uni(b.sp_A[P]), C = b.sp_A[Qp], Dsh = b.sp_B[Qp];
    long long N, msh;
    mol_layout(b, A, N, msh);
    Jo = shifted(Jo, msh);
    const int aoC = b.shell_ao[C] + CK::c0, aoD = b.shell_ao[Dsh];
#pragma unroll
    for (int c = 0; c < NCDL; ++c) acc_add(&Jo[(aoC + c / ND) * N + aoD + c % ND], jcd[c], FX);
  }
}

// Wave item `item` of a unit: a C chunk with its A chunks for CARRY plans,
// one chunk otherwise.
template <class PL, int CC, typename CT, bool FX>
__device__ __forceinline__ void dispatch_citem(int cc, const DevBasis& b, const EriOut& o, int P, int Qp, int pos,
                                               int nab, bool active, const CT* __restrict__ tab, const RowAcc& acc,
                                               long long jb) {
  if constexpr (!PL::CARRY) {
    dispatch_chunk<PL, 0, CT, FX>(cc, b, o, P, Qp, pos, nab, active, tab, acc, jb);
  } else if constexpr (CC < PL::SPLIT_C) {
    if (cc == CC) {
      run_citem<PL, CC, CT, FX>(b, o, P, Qp, pos, nab, active, tab, acc, jb);
      return;
    }
    dispatch_citem<PL, CC + 1, CT, FX>(cc, b, o, P, Qp, pos, nab, active, tab, acc, jb);
  }
}

// Work mapping: each workgroup pulls row tasks (wave units of one bra pair,
// longest first) from the class pair's queue; its eight wavefronts take the task's
// (unit, chunk) items in turn, so within a wave the bra pair and the chunk
// (and with it all register indexing) are uniform. With LDS accumulators the
// task's J[a,b] block and K rows of a and b are summed in LDS and added to
// J/K once at the end of the task.
//
// Precision: CT is the kernel's arithmetic. Without o.mixed only the FP64
// kernel runs and takes every unit. With o.mixed both run, each walking the
// whole task list from its own queue head: a unit goes to the FP32 kernel when
// every active quartet's bound Q_ab Q_cd max|D| is below tau (a wave-uniform
// vote), else to the FP64 kernel, and the other kernel skips it. Separate
// kernels keep the FP32 path's register allocation (and so its occupancy)
// its own.
//
// The read-only inputs are also passed as __restrict__ arguments and rebound
// into b, w and o: with no possible aliasing store (the J/K atomics) the
// compiler can turn wave-uniform loads (wave unit, bra pair, bra primitive
// pairs, D blocks of the bra) into scalar loads.
// FX: accumulate in the 64-bit fixed point of deterministic builds
// (EriOut::fixed), a separate instantiation so that the FP64-atomic kernels
// carry none of its code or registers.
template <class PL, typename CT, bool FX>
__global__ __launch_bounds__(kGroupThreads) void eri_fast_kernel(DevBasis b, ClassWork w, EriOut o,
                                                       const PrimPairF* __restrict__ ppf,
                                                       const PrimPairF32* __restrict__ ppf32,
                                                       const WaveUnit* __restrict__ units,
                                                       const int* __restrict__ col_pair,
                                                       const int* __restrict__ sp_A, const int* __restrict__ sp_B,
                                                       const int* __restrict__ pp_start,
                                                       const int* __restrict__ pp_count,
                                                       const double* __restrict__ shell_xyz,
                                                       const int* __restrict__ shell_ao,
                                                       const double* __restrict__ Qb,
                                                       const double* __restrict__ Dsp,
                                                       const double* __restrict__ Dm) {
  b.ppf = ppf;
  b.ppf32 = ppf32;
  b.sp_A = sp_A;
  b.sp_B = sp_B;
  b.pp_start = pp_start;
  b.pp_count = pp_count;
  b.shell_xyz = shell_xyz;
  b.shell_ao = shell_ao;
  b.Q = Qb;
  b.Dsp = Dsp;
  w.units = units;
  w.col_pair = col_pair;
  o.D = Dm;
  constexpr bool kF32 = sizeof(CT) == 4;
  constexpr int NIT = PL::CARRY ? PL::SPLIT_C : PL::NCH, NR = PL::NR;  // wave items per unit
  constexpr int NA = nc_(PL::la), NB = PL::NB;
  constexpr bool kLds = NR <= kLdsRootsMax;
  // LDS copy of the CT root table, intervals padded (rys_c::lds_stride_of)
  constexpr int NFI = rys_c::interval_size_of<CT>(NR);
  constexpr int SCT = rys_c::lds_stride_of<CT>(NR);
  constexpr int TB = kLds ? rys_c::block_size_of<CT>(NR) : 1;
  constexpr int TBC = kLds ? rys_c::nint_of<CT>(NR) * SCT : 1;
  const int wpg = w.wpg, nthr = 64 * w.wpg;  // the launch's block size (uniform)
  __shared__ CT tct[TBC];
  const CT* gct;
  if constexpr (kF32)
    gct = rys_dev_f::rys32_coef + rys_c::coef_off32[NR];
  else
    gct = rys_dev::rys_coef + rys_c::coef_off[NR];
  if constexpr (kLds) {
    for (int t = threadIdx.x; t < TB; t += nthr) {
      const int iv = t / NFI, f = t - iv * NFI;
      tct[iv * SCT + f] = gct[t];
    }
  }
  const bool use_acc = w.acc_colsC > 0;
  const int nden = o.D2 ? 2 : 1;  // densities digested (accumulator sets in s_acc)
  RowAcc acc{};
  int nacc = 0;
  if (use_acc) {  // row-task accumulators in s_acc, sized at launch
    const int cC = w.acc_colsC, cD = w.acc_colsD;  // cD == 0: D columns share the C space
    acc.on = true;
    acc.colsC = cC;
    acc.colsD = cD > 0 ? cD : cC;
    acc.jab = 0;
    acc.kac = NA * NB;
    acc.kad = cD > 0 ? acc.kac + NA * cC : acc.kac;
    acc.kbc = acc.kac + NA * cC + (cD > 0 ? NA * cD : 0);
    acc.kbd = cD > 0 ? acc.kbc + NB * cC : acc.kbc;
    nacc = NA * NB + (NA + NB) * (cC + cD);
    acc.stride = nacc;
    for (int t = threadIdx.x; t < nacc * nden; t += nthr) s_acc[t] = 0.0;
  }
  __shared__ int s_task, s_work;
  if (threadIdx.x == 0) s_work = 0;
  __syncthreads();
  const CT* tab = kLds ? tct : gct;
  const int lane = threadIdx.x & 63;

  unsigned int* queue = kF32 ? w.queue32 : w.queue;
  // launched with kGroupThreads-thread blocks; blockDim is not read because
  // the compiler cannot prove it uniform, which would make the loops divergent
  const int wib = __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
  for (;;) {  // persistent workgroups pull row tasks from the class pair's queue
    if (threadIdx.x == 0) s_task = (int)atomicAdd(queue, 1u);
    __syncthreads();
    const int t = s_task;
    if (t >= w.ntasks) break;  // uniform per workgroup; every workgroup gets here
    const RowTask tk = w.tasks[t];  // routed lists hold empty tasks (nu = 0): no unit to read
    int P = 0;
    if (tk.nu > 0) P = uni(w.units[tk.u0].P);
    bool worked = false;  // this wave digested an item of the task (uniform)
    // a wave item is one chunk of a unit, or for CARRY plans one C chunk with
    // its A chunks in sequence; items chunk-major: the waves of a workgroup
    // run the same chunk's code
    // over different units (a split plan's chunks are separate code regions,
    // together far larger than the instruction cache)
    for (int item = wib; item < tk.nu * NIT; item += wpg) {  // uniform per wave
      const int chunk = item / tk.nu;
      const int ui = item - chunk * tk.nu;
      const WaveUnit un = w.units[tk.u0 + ui];
      bool active = lane < un.ncol;
      // the lane's ket position: units start at 64-position block boundaries,
      // so every lane's position (tail lanes: padding) is inside the lists
      const int pos = un.col + lane;
      const int Qp = w.col_pair[active ? pos : un.col];
      double bound = 0.0;
      if (active) {
        bound = (b.Q[P] * b.kq[pos]) * fmax(b.Dsp[P], b.kd[pos]);
        if (o.eps_density > 0.0) active = bound >= o.eps_density;
      }
      if (!wave_any(active)) continue;  // the whole unit is density-screened
      if (!w.routed) {  // without routed lists each kernel takes the units its vote gives it
        const bool lowp = o.mixed && wave_all(!active || bound < o.tau_mixed);
        if (lowp != kF32) continue;  // the other precision's kernel takes this unit
      }
      worked = true;
      if (o.counters && chunk == 0) {
        // statistics: one atomic per counter and wave
        const int A = b.sp_A[P], B = b.sp_B[P], C = b.sp_A[Qp], Dsh = b.sp_B[Qp];
        unsigned long long* cnt = o.counters + kCountStride * w.cp_index;
        const unsigned long long kq =
            active ? (unsigned long long)b.pp_count0[P] * (unsigned long long)b.pp_count0[Qp] : 0ull;
        const int bin = (((A != B) << 2) | ((C != Dsh) << 1) | (P != Qp));
        for (int k = 0; k < 8; ++k) {
          const unsigned long long m = __ballot(active && bin == k);
          if (m && lane == 0) atomicAdd(cnt + k, (unsigned long long)__popcll(m));
        }
        const unsigned long long ksum = (unsigned long long)wave_sum((double)kq);
        const unsigned long long nact = (unsigned long long)__popcll(__ballot(active));
        // primitive quartets the loops evaluate (after the primitive-quartet
        // cut) and the lane slots the wave spends on them (64 x the longest lane)
        const int nabc = uni(o.eps_prim > 0.0 ? un.nab : b.pp_count[P]);
        long long nev = 0;
        if (active && nabc > 0) {
          const int cd0 = b.pp_start[Qp], ncd = b.pp_count[Qp];
          const double kcut = o.eps_prim > 0.0 ? o.eps_prim / b.ppf[b.pp_start[P]].q : 0.0;
          int j = 0;
          while (j < ncd && !(b.ppf[cd0 + j].q < kcut)) ++j;
          nev = (long long)nabc * j;
        }
        const double nevs = wave_sum((double)nev);
        double nmax = (double)nev;
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) nmax = fmax(nmax, __shfl_xor(nmax, off));
        if (lane == 0) {
          atomicAdd(cnt + 8, ksum);
          atomicAdd(cnt + 11, (unsigned long long)nevs);
          atomicAdd(cnt + 12, 64ull * (unsigned long long)nmax);
          if (kF32) {
            atomicAdd(cnt + 9, nact);
            atomicAdd(cnt + 10, ksum);
          }
        }
      }
      // inactive lanes (tail, screened) run with an empty primitive loop and no
      // atomics so the wave stays converged for the cross-lane reductions
      // bra primitive pairs this unit evaluates (see run_chunk)
      const int nab = uni(o.eps_prim > 0.0 ? un.nab : b.pp_count[P]);
      // the unit's block of the ket-position buffers (units start at block boundaries)
      const long long jb = o.jblk ? o.jblk[un.col >> 6] : -1LL;
      dispatch_citem<PL, 0, CT, FX>(chunk, b, o, P, Qp, pos, nab, active, tab, acc, jb);
    }
    if (use_acc) {
      if (worked && lane == 0) s_work = 1;
      __syncthreads();  // the task's contributions are in LDS
      if (s_work) {
        const int A = b.sp_A[P], B = b.sp_B[P];
        long long N, msh;
        mol_layout(b, A, N, msh);
        const long long aoA = b.shell_ao[A], aoB = b.shell_ao[B];
        const int cC = w.acc_colsC, cD = w.acc_colsD;
        const int* mlao = b.mol_lao + b.shell_mol[A] * (ASUKA_LMAX + 1);  // the molecule's AO lists
        const int* laoC = b.lao_ao + w.lao_offC + mlao[PL::lc];
        const int* laoD = b.lao_ao + (cD > 0 ? w.lao_offD + mlao[PL::ld] : w.lao_offC + mlao[PL::lc]);
        const int cDe = cD > 0 ? cD : cC;
        // compact columns of the C and D lists (EriOut::Kc)
        const int gC = w.lao_offC + mlao[PL::lc], gD = cD > 0 ? w.lao_offD + mlao[PL::ld] : gC;
        for (int t2 = threadIdx.x; t2 < nacc * nden; t2 += nthr) {
          const double v = s_acc[t2];  // fixed builds: the bits of an integer sum
          if (__double_as_longlong(v) == 0LL) continue;
          s_acc[t2] = 0.0;
          const bool second = t2 >= nacc;
          double* Jo = shifted(second ? o.J2 : ((o.mode == MODE_FOCK) ? o.F : o.J), msh);
          double* Ko = shifted(second ? o.K2 : ((o.mode == MODE_FOCK) ? o.F : o.K), msh);
          double* Kc = second ? o.Kc2 : o.Kc;  // uniform
          int r = second ? t2 - nacc : t2;
          if (r < NA * NB) {
            acc_add_sum(&Jo[(aoA + r / NB) * N + aoB + r % NB], v, FX);
            continue;
          }
          r -= NA * NB;
          long long row, k;  // the K row (AO) and the entry's index in its column list
          int g;             // first compact column of that list
          const int* lao;
          if (r < NA * cC) {
            row = aoA + r / cC, k = r % cC, g = gC, lao = laoC;
          } else if ((r -= NA * cC), cD > 0 && r < NA * cD) {
            row = aoA + r / cD, k = r % cD, g = gD, lao = laoD;
          } else {
            if (cD > 0) r -= NA * cD;
            if (r < NB * cC) {
              row = aoB + r / cC, k = r % cC, g = gC, lao = laoC;
            } else {
              r -= NB * cC;
              row = aoB + r / cDe, k = r % cDe, g = gD, lao = laoD;
            }
          }
          if (Kc)  // consecutive entries, consecutive addresses
            acc_add_sum(&Kc[row * N + g + k], v, FX);
          else
            acc_add_sum(&Ko[row * N + lao[k]], v, FX);
        }
      }
    }
    __syncthreads();  // accumulators are zero again; s_task and s_work may be overwritten
    if (use_acc && threadIdx.x == 0) s_work = 0;
  }
}

// Bytes of LDS accumulators for a class pair with ket columns (colsC, colsD),
// per density.
inline size_t row_acc_bytes(int la, int lb, int colsC, int colsD) {
  const int NA = nc_(la), NB = nc_(lb);
  return sizeof(double) * (size_t)(NA * NB + (NA + NB) * (colsC + colsD));
}

// The FP32 kernel's plan of a class: half the FP64 plan's component chunks
// (the FP32 intermediates take half the registers), A chunks halved when
// their count is even, else C chunks; ASUKA_F32_HALVE=0 keeps the FP64 plan's.
// Measured at (H2O)64 cc-pVDZ, single stream: (21|11) 9.6 -> 7.6 ms, (20|11)
// 28.3 -> 18.3 ms, (21|21) 18.0 -> 11.7 ms per build for the FP32 kernels.
#ifndef ASUKA_F32_HALVE
#define ASUKA_F32_HALVE 1
#endif
template <class PL>
struct F32PlanOf {
  static constexpr bool halve_a = ASUKA_F32_HALVE && PL::SPLIT_A % 2 == 0;
  static constexpr bool halve_c = ASUKA_F32_HALVE && !halve_a && PL::SPLIT_C % 2 == 0;
  static constexpr int SA = halve_a ? PL::SPLIT_A / 2 : PL::SPLIT_A;
  static constexpr int SC = halve_c ? PL::SPLIT_C / 2 : PL::SPLIT_C;
  using type = Plan<PL::la, PL::lb, PL::lc, PL::ld, SA, SC>;
};

template <class PL, typename CT, bool FX = false>
hipError_t launch_one(const DevBasis& b, const ClassWork& w, const EriOut& o, size_t lds, long long g, hipStream_t s) {
  if (lds > 64 * 1024) {
    const hipError_t e = hipFuncSetAttribute((const void*)eri_fast_kernel<PL, CT, FX>,
                                             hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    if (e != hipSuccess) return e;
  }
  hipLaunchKernelGGL((eri_fast_kernel<PL, CT, FX>), dim3((unsigned)g), dim3(64 * w.wpg), lds, s, b, w, o, b.ppf, b.ppf32,
                     w.units, w.col_pair, b.sp_A, b.sp_B, b.pp_start, b.pp_count, b.shell_xyz, b.shell_ao, b.Q,
                     b.Dsp, o.D);
  return hipGetLastError();
}

// The FP64 kernel of a class pair and, for mixed builds, the FP32 kernel after
// it on the same stream (each takes its own units, see eri_fast_kernel).
template <class PL>
hipError_t launch_plan(const DevBasis& b, const ClassWork& w, const EriOut& o, hipStream_t s) {
  long long g = w.ntasks;  // persistent: a few workgroups per CU, tasks from the queue
  const long long gmax = w.wpg == 8 ? 1024 : 2048;
  if (g > gmax) g = gmax;
  if (g < 1) g = 1;
  const size_t lds =
      w.acc_colsC > 0 ? (o.D2 ? 2 : 1) * row_acc_bytes(PL::la, PL::lb, w.acc_colsC, w.acc_colsD) : 0;
  // deterministic builds (no mixed routing: every unit on the FP64 kernel)
  if (o.fixed) return launch_one<PL, double, true>(b, w, o, lds, g, s);
  hipError_t e = launch_one<PL, double>(b, w, o, lds, g, s);
  if (e != hipSuccess || !o.mixed) return e;
  ClassWork w32 = w;
  if (w.routed) {
    w32.units = w.units32;
    w32.tasks = w.tasks32;
  }
  return launch_one<typename F32PlanOf<PL>::type, float>(b, w32, o, lds, g, s);
}

using Launcher = hipError_t (*)(const DevBasis&, const ClassWork&, const EriOut&, hipStream_t);

}  // namespace asuka_fast

// Each jk_fast_*.hip translation unit instantiates a subset of class plans;
// FN returns a plan's launcher, FN##_nch its wave items per unit (the number
// of component chunks, or of C chunks for plans that carry J[c,d] across
// their A chunks).
#define ASUKA_FAST_FIND(FN, LIST)                                                     \
  asuka_fast::Launcher FN(int la, int lb, int lc, int ld) {                          \
    using namespace asuka_fast;                                                       \
    LIST(ASUKA_FAST_CASE)                                                             \
    return nullptr;                                                                   \
  }                                                                                   \
  int FN##_nch(int la, int lb, int lc, int ld) {                                      \
    LIST(ASUKA_FAST_NCH_CASE)                                                         \
    return 0;                                                                         \
  }
#define ASUKA_FAST_CASE(a, b_, c, d, sa, sc) \
  if (la == a && lb == b_ && lc == c && ld == d) return &launch_plan<Plan<a, b_, c, d, sa, sc>>;
#define ASUKA_FAST_NCH_CASE(a, b_, c, d, sa, sc) \
  if (la == a && lb == b_ && lc == c && ld == d) return asuka_fast::Plan<a, b_, c, d, sa, sc>::CARRY ? (sc) : (sa) * (sc);
