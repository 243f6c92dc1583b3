// Internal types shared by the J/K translation units (not part of the ABI).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <string>
#include <vector>

#define ASUKA_LMAX 3  // highest shell angular momentum of this build (f)
#define ASUKA_NCART_MAX 10
#define ASUKA_NROOTS_MAX 7

// The register kernels copy the root tables of up to this many roots into
// LDS (each workgroup); larger root counts read them from global memory.
#ifndef ASUKA_LDS_ROOTS_MAX
#define ASUKA_LDS_ROOTS_MAX 4
#endif
constexpr int kLdsRootsMax = ASUKA_LDS_ROOTS_MAX;

// One primitive pair of a shell pair (only pairs with a nonzero coefficient
// product are kept; dropping exact zeros leaves every integral unchanged).
struct __align__(16) PrimPair {
  double eta;     // a + b
  double Px, Py, Pz;
  double cK;      // ca * cb * exp(-a b / eta |AB|^2)
  double a_exp;   // exponent of the bra-side primitive (A)
  double b_exp;   // exponent of the ket-side primitive (B)
  double inv_eta;
};

// The same primitive pair in the form the class-specialised kernels read:
// everything the primitive loop needs, nothing else. ck folds the Rys
// prefactor: the product ck_ab ck_cd / sqrt(eta_ab + eta_cd) is
// 2 pi^(5/2) cK_ab cK_cd / (eta_ab eta_cd sqrt(eta_ab + eta_cd)).
struct __align__(16) PrimPairF {
  double eta, inv_eta, ck;
  double Px, Py, Pz;
  double PAx, PAy, PAz;  // P minus the centre of the pair's first shell
  double q;              // Schwarz bound of the primitive pair alone (pairs sorted by it, descending)
};

// The primitive pair as the FP32 path of the class-specialised kernels reads
// it: the same quantities rounded to float, except the centre P, kept in FP64
// so that P - Q is formed in FP64 before rounding.
struct __align__(16) PrimPairF32 {
  double Px, Py, Pz;
  float eta, inv_eta, ck;
  float PAx, PAy, PAz;
  float q;
  float pad;
};

enum EriMode : int {
  MODE_JK = 0,       // accumulate J and/or K
  MODE_FOCK = 1,     // accumulate J - K/2 into F
  MODE_TILE = 2,     // write the tile
  MODE_SCHWARZ = 3,  // Q[P] = sqrt(max diag (P|P))
  MODE_X3C = 4,      // three-centre (ab|P): X[P][a][b] (and [b][a]), ket pair = (aux shell, unit s)
  MODE_V2C = 5,      // two-centre (P|Q): V[P][Q], bra and ket pairs = (aux shell, unit s)
};

// Device-side view of a context (passed by value to kernels).
struct DevBasis {
  const double* shell_xyz;   // (nshell, 3)
  const int* shell_l;
  const int* shell_ao;
  const int* sp_A;
  const int* sp_B;
  const int* pp_start;       // (nsp) first primitive pair of the shell pair
  const int* pp_count;       // (nsp)
  const int* pp_count0;      // (nsp) primitive pairs before primitive-pair screening (work statistics)
  const PrimPair* pp;
  const PrimPairF* ppf;      // same indexing as pp
  const PrimPairF32* ppf32;  // same indexing as pp
  const int* shell_lao;      // (nshell) first AO of the shell in the compact numbering of its l
  const int* lao_ao;         // compact (l, index) -> AO, concatenated over l
  const int* shell_mol;      // (nshell) molecule of the shell (0 unless a batch context)
  const int* mol_lao;        // (nmol, LMAX+1) start of a molecule's AO list within its l part of lao_ao
  const int* mol_n;          // (nmol) row stride of the molecule's D/J/K block
  const long long* mol_shift;  // (nmol) element offset of that block (see mol_layout)
  const double* Q;           // (nsp) Schwarz bounds
  const double* Dsp;         // (nsp) max |D| per shell-pair block
  // Per ket position (the padded ket lists of the wave units, EriOut::jblk),
  // read with contiguous per-lane loads instead of gathers through the pair
  // index: the pair's Q and max |D| (kd, written by each build's k_dsp), its
  // first primitive pair and count, the first AOs of its shells C and D,
  // their compact (per-l) columns, and C - D (3 doubles per position).
  const double* kq;
  const double* kd;
  const int2* kpp;
  const int2* kao;
  const int2* klao;
  const double* kcd;
  // The ket primitive pairs by position, transposed so that the lanes of a
  // wave read the j-th primitive pair of their kets at consecutive addresses:
  // block k's primitive j of lane l is slot (kpoff[k] + j) * 64 + l, field f
  // of slot t at kp64[f * kslots + t] (PrimPairF's ten fields in order, then
  // the exponent of the pair's first primitive (PrimPair::a_exp, for the
  // derivative kernels); FP64 kernels) and, for the FP32 kernels,
  // kp32d[f * kslots + t] (Px, Py, Pz) and kp32f[f * kslots + t] (eta,
  // inv_eta, ck, PAx, PAy, PAz, q).
  const long long* kpoff;
  const double* kp64;
  const double* kp32d;
  const float* kp32f;
  long long kslots;
  int nao;
  int nsp;
};

// Work of one class pair. Quartets are grouped into segments by bra and ket
// contraction depth (primitive pairs per shell pair, in buckets at most 25 %
// wide), so the quartets of a segment have nearly equal primitive-loop
// lengths. In a segment, rows are bra pairs
// (Q-descending) and each row runs over a prefix of the segment's ket list
// (Q-descending). Segment starts are padded to multiples of 64 in the class
// pair's work index space, so a wavefront never spans two segments; padding
// indices decode as invalid.
struct Segment {
  long long off;   // first work index (multiple of 64)
  long long len;   // quartets in the segment
  int row_start;   // into row_pair / row_off (row_off has nrows + 1 entries)
  int nrows;
  int col_start;   // into col_pair
  int kab, kcd;    // largest primitive-pair count of the segment's bra / ket pairs (bucket)
  int pad;
};

// One wavefront's share of a segment: a bra pair and up to 64 consecutive
// kets of its row (one per lane), so the bra side is uniform across the wave.
struct WaveUnit {
  int P;     // bra shell pair
  int col;   // first ket, index into col_pair
  int ncol;  // 1..64
  int nab;   // bra primitive pairs to evaluate: those whose bound times the
             // largest ket-pair bound of the unit reaches the primitive-quartet cut
};

// A workgroup's share: consecutive wave units [u0, u0 + nu) of one bra row.
struct RowTask {
  int u0, nu;
};

struct ClassWork {
  int la, lb, lc, ld;
  const WaveUnit* units;  // class-specialised kernels (bra-uniform wavefronts)
  long long nunits;
  const RowTask* tasks;
  long long ntasks;
  // LDS row accumulators of the ket class's columns: the number of AOs of
  // angular momentum lc (acc_colsC) and ld (acc_colsD, 0 when lc == ld) and
  // where their AO lists start in lao_ao; acc_colsC == 0 disables them.
  int acc_colsC, acc_colsD, lao_offC, lao_offD;
  unsigned int* queue;  // work-queue head (row tasks handed out in order), zero at launch
  unsigned int* queue32;  // the FP32 register kernel's own head (mixed builds), zero at launch
  int wpg;              // wavefronts per workgroup of the register kernels (4 or 8)
  const Segment* seg;
  int nseg;
  const int* row_pair;
  const long long* row_off;
  const int* col_pair;
  long long nwork;  // padded work index range
  // explicit quartet list (tile test hook); when set, segments are unused
  const int* q_ab;
  const int* q_cd;
  const long long* q_out_off;
  // product list (three- and two-centre integrals): work index k is the
  // quartet (q_bra0 + k / q_nket, q_ket0 + k % q_nket) when q_nket > 0
  long long q_nket;
  int q_bra0, q_ket0;
  int cp_index;  // class-pair index (statistics counters)
  // mixed builds with routed units (the vote taken once per build before the
  // launches): the FP32 kernel's units and row tasks; units and tasks above
  // then hold the FP64 kernel's. routed == 0: each kernel votes in place.
  const WaveUnit* units32;
  const RowTask* tasks32;
  int routed;
};

struct EriOut {
  int mode;
  const double* D;
  double* J;  // MODE_JK (may be null)
  double* K;  // MODE_JK (may be null)
  double* F;  // MODE_FOCK
  double* tile;
  double* Qout;
  double eps_density;  // <= 0 disables density screening
  // optional statistics, kCountStride per class pair: [0..7] quartets evaluated
  // by symmetry bin (A!=B)<<2 | (C!=D)<<1 | (bra pair != ket pair), [8] their
  // primitive quartets, [9] quartets evaluated in FP32, [10] their primitive quartets
  unsigned long long* counters;
  double eps_prim;    // primitive quartets with Q_p Q_q below this are skipped (0: none)
  int mixed;          // allow the FP32 path
  double tau_mixed;   // FP32 when Q_ab Q_cd max|D| < tau_mixed for the whole wave
  // MODE_JK with two densities (UHF/ROHF): each quartet is evaluated once and
  // digested with D into (J, K) and with D2 into (J2, K2); D2 == nullptr for one
  const double* D2;
  double* J2;
  double* K2;
  // MODE_X3C: X (x_naux, x_nao, x_nao); MODE_V2C: X (x_naux, x_naux); AO and
  // auxiliary function indices are the shells' shell_ao
  double* X;
  long long x_nao, x_naux;
  // Ket-position buffers of the register kernels (null: J[c,d] by atomics on
  // J and D[c,d] read from D). Ket lists (col_pair) are laid out in blocks of
  // 64 positions of one ket class; block k holds component c (= ic * ND + id)
  // of the pair at position 64 k + lane at jblk[k] + 64 c + lane. A wave
  // unit starts at a block boundary, so its lanes read D[c,d] (Dpos, gathered
  // before the launches) and add J[c,d] (Jpos, added into J after them) with
  // contiguous addresses. Jpos2 / Dpos2: the second density's.
  const long long* jblk;
  double* Jpos;
  double* Jpos2;
  const double* Dpos;
  const double* Dpos2;
  // K rows of the register kernels' row-task flush with permuted columns
  // (one-molecule contexts; null: into K itself): column g holds AO
  // lao_ao[g], so a row's columns of one angular momentum are contiguous
  // and the flush's atomics coalesce; added into K (or F) after the
  // launches. Kc2: the second density's.
  double* Kc;
  double* Kc2;
  // ASUKA_JK_DETERMINISTIC: J, K, F, Jpos and Kc hold 64-bit two's-complement
  // fixed point (units of 2^-44) during the build and every contribution is
  // added with an integer atomic (acc_add), so the result does not depend on
  // the order of the adds
  int fixed;
};

// Fixed-point unit of deterministic builds: 2^44 per unit value (resolution
// 5.7e-14, range +-2^19 for the 64-bit sums).
constexpr double kFixScale = 17592186044416.0;

// p += v: an FP64 atomic add, or with fixed set an integer atomic add of v in
// units of 2^-44 (p then holds the 64-bit fixed-point sum).
__device__ __forceinline__ void acc_add(double* p, double v, bool fixed) {
  if (fixed)
    atomicAdd(reinterpret_cast<unsigned long long*>(p), (unsigned long long)__double2ll_rn(v * kFixScale));
  else
    atomicAdd(p, v);
}
// p += a partial sum read from an accumulator of the same representation
// (fixed: the 64-bit pattern of an integer sum).
__device__ __forceinline__ void acc_add_sum(double* p, double v, bool fixed) {
  if (fixed)
    atomicAdd(reinterpret_cast<unsigned long long*>(p), (unsigned long long)__double_as_longlong(v));
  else
    atomicAdd(p, v);
}

constexpr int kCountStride = 16;

// D/J/K layout: element (a, b) of a matrix X, for AOs a, b (global indices) of
// the molecule that shell A belongs to, is X[shift + a * stride + b]. One
// molecule: stride nao, shift 0 (plain row-major nao x nao). A batch context
// packs each molecule's nao_m x nao_m block at its own base:
// stride nao_m, shift base_m - off_m (nao_m + 1), off_m its first AO.
__device__ inline void mol_layout(const DevBasis& b, int A, long long& stride, long long& shift) {
  const int m = b.shell_mol[A];
  stride = b.mol_n[m];
  shift = b.mol_shift[m];
}
template <typename T>
__device__ inline T* shifted(T* p, long long sh) {
  return p ? p + sh : p;
}

__host__ __device__ constexpr inline int ncart_c(int l) { return (l + 1) * (l + 2) / 2; }

// v of the lane that DPP control CTRL selects (both 32-bit halves moved);
// lanes of rows outside ROWS receive 0.
template <int CTRL, int ROWS = 0xF>
__device__ __forceinline__ double dpp_f64(double v) {
  const unsigned long long x = (unsigned long long)__double_as_longlong(v);
  int lo, hi;
  if constexpr (ROWS == 0xF) {  // every row written: no old value to keep
    lo = __builtin_amdgcn_mov_dpp((int)(unsigned)x, CTRL, ROWS, 0xF, false);
    hi = __builtin_amdgcn_mov_dpp((int)(unsigned)(x >> 32), CTRL, ROWS, 0xF, false);
  } else {
    lo = __builtin_amdgcn_update_dpp(0, (int)(unsigned)x, CTRL, ROWS, 0xF, false);
    hi = __builtin_amdgcn_update_dpp(0, (int)(unsigned)(x >> 32), CTRL, ROWS, 0xF, false);
  }
  return __longlong_as_double((long long)(((unsigned long long)(unsigned)hi << 32) | (unsigned)lo));
}

// Sum of v over the 64 lanes of the wavefront, the same on every lane; all
// lanes must be enabled. The lanes are summed in registers with DPP moves, not
// through the LDS crossbar (ds_bpermute): within each 16-lane row by xor-1 and
// xor-2 quad permutes and the half-row and row mirrors, then across rows by
// the row 15 and row 31 broadcasts, which leave the total on lane 63.
__device__ __forceinline__ double wave_total(double v) {
  v += dpp_f64<0xB1>(v);   // quad_perm [1,0,3,2]
  v += dpp_f64<0x4E>(v);   // quad_perm [2,3,0,1]
  v += dpp_f64<0x141>(v);  // row_half_mirror: lane i <- 7 - i within 8
  v += dpp_f64<0x140>(v);  // row_mirror: lane i <- 15 - i within 16
  v += dpp_f64<0x142, 0xA>(v);  // row_bcast:15 into rows 1 and 3
  v += dpp_f64<0x143, 0xC>(v);  // row_bcast:31 into rows 2 and 3
  const unsigned long long x = (unsigned long long)__double_as_longlong(v);
  const unsigned lo = (unsigned)__builtin_amdgcn_readlane((int)(unsigned)x, 63);
  const unsigned hi = (unsigned)__builtin_amdgcn_readlane((int)(unsigned)(x >> 32), 63);
  return __longlong_as_double((long long)(((unsigned long long)hi << 32) | lo));
}

// Decode work index k of a class pair into (bra pair, ket pair); false for
// segment padding.
__device__ inline bool decode_work(const ClassWork& w, long long k, int& P, int& Qp) {
  if (w.q_nket > 0) {
    P = w.q_bra0 + (int)(k / w.q_nket);
    Qp = w.q_ket0 + (int)(k % w.q_nket);
    return true;
  }
  if (w.q_ab) {
    P = w.q_ab[k];
    Qp = w.q_cd[k];
    return true;
  }
  int lo = 0, hi = w.nseg - 1;
  while (lo < hi) {  // largest segment with off <= k
    const int mid = (lo + hi + 1) >> 1;
    if (w.seg[mid].off <= k)
      lo = mid;
    else
      hi = mid - 1;
  }
  const Segment sg = w.seg[lo];
  const long long kk = k - sg.off;
  if (kk >= sg.len) return false;
  const long long* ro = w.row_off + sg.row_start;
  int a = 0, c = sg.nrows - 1;
  while (a < c) {  // largest row with ro[r] <= kk
    const int mid = (a + c + 1) >> 1;
    if (ro[mid] <= kk)
      a = mid;
    else
      c = mid - 1;
  }
  P = w.row_pair[sg.row_start + a];
  Qp = w.col_pair[sg.col_start + (int)(kk - ro[a])];
  return true;
}

// Launchers (defined in jk_generic.hip / jk_fast.hip).
hipError_t launch_eri_generic(const DevBasis& b, const ClassWork& w, const EriOut& o, hipStream_t s);
// Wave-cooperative class templates (jk_coop.hip); *handled is false for classes it does not serve.
hipError_t launch_eri_coop(const DevBasis& b, const ClassWork& w, const EriOut& o, hipStream_t s, bool* handled);
