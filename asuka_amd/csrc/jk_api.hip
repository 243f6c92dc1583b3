// C-ABI of the J/K engine: context construction, builds, test hooks.
//
// Context construction mirrors make_direct_jk_context (asuka/hf/direct_jk.py:1143-1470):
//   shell pairs with l_A >= l_B ................ asuka/cueri/shell_pairs.py:72-116
//   primitive-pair tables (eta, P, cK) ......... asuka/cueri/_eri_rys_cpu.pyx:2066-2129
//   Schwarz bounds on the device ............... asuka/cueri/screening.py:224-314
//   Q-sorted, screened quartet rows ............ asuka/hf/direct_jk.py:1228-1246
// but instead of materialising 8 B/quartet task arrays grouped by class, the
// quartets of each (bra class, ket class) pair are described by rows: a bra
// pair and the length of the Q-descending prefix of the ket class list it
// pairs with. Kernels decode quartets from (row, column) on the fly.
#include <hip/hip_runtime.h>
#include <hipcub/hipcub.hpp>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <memory>
#include <mutex>
#include <numeric>
#include <string>
#include <vector>

#include "../../include/asuka_jk.h"
#include "jk_internal.h"
#include "rys_eval.h"

hipError_t launch_eri_fast(const DevBasis& b, const ClassWork& w, const EriOut& o, hipStream_t s, bool* handled);
bool eri_fast_has_plan(int la, int lb, int lc, int ld);
int eri_fast_nch(int la, int lb, int lc, int ld);

static thread_local std::string g_err;

// default FP32 threshold on Q_ab Q_cd max|D| for builds with ASUKA_JK_MIXED
static constexpr double kDefaultTauMixed = 1e-5;
// Row tasks: at most this many wave units of one bra row per workgroup.
// Long tasks flush the LDS row accumulators less often: measured at (H2O)64
// cc-pVDZ, at least 128 units per task took the FP64 Fock build from 476 to
// 447 ms and the mixed one from 396 to 332 ms against 16-32, and left (H2O)8
// cc-pVTZ unchanged (a task never spans two bra rows, so short rows stay short).
static constexpr int kTaskUnits = 128;
// LDS accumulators of one row task larger than this (or than the device allows
// next to the root tables) fall back to global atomics; ASUKA_ROW_ACC_MAX_KB
// lowers it. gfx950 gives a workgroup up to 160 KB: accumulators above 80 KB
// leave one workgroup per CU (8 wavefronts), which the register kernels of
// the classes that need them (2 wavefronts per SIMD by registers) fill anyway.
static constexpr size_t kRowAccMaxBytes = 160 * 1024;
// streams a build's class-pair launches are spread over (the caller's included)
static constexpr int kDefaultStreams = 4;
static constexpr int kMaxStreams = 8;
// LDS the register kernels of a class with nr roots hold for root tables
static inline size_t root_table_lds(int nr) { return nr <= kLdsRootsMax ? rys_c::lds_table_bytes(nr) : 0; }
// Accumulators up to this size run with 4-wave workgroups (16-unit tasks),
// larger ones with 8-wave workgroups (32-unit tasks).
static constexpr size_t kRowAccSmallBytes = 40 * 1024;
// sqrt(2 pi^(5/2)), split evenly over the bra and ket primitive-pair factors
static constexpr double kSqrt2Pi52 = 5.91496717279561287782;
// primitive pairs with Q_p * max Q_p below this are dropped (with eps_schwarz > 0)
static constexpr double kPrimPairEps = 1e-17;
// primitive quartets with Q_p Q_q below this are skipped by the register
// kernels (with eps_schwarz > 0)
static constexpr double kPrimQuartetEps = 1e-17;

static int fail(int code, const std::string& msg) {
  g_err = msg;
  return code;
}

// the same for the other translation units of the library
int asuka_fail_msg(int code, const char* msg) { return fail(code, msg ? msg : ""); }

#define HIP_CHECK(expr)                                                                       \
  do {                                                                                        \
    hipError_t _e = (expr);                                                                   \
    if (_e != hipSuccess) return fail(ASUKA_E_HIP, std::string(#expr) + ": " + hipGetErrorString(_e)); \
  } while (0)

namespace {

int pair_class(int la, int lb) { return la * (la + 1) / 2 + lb; }  // la >= lb

struct ClassPairHost {
  int X, Y;  // bra class >= ket class
  int la, lb, lc, ld;
  int seg_start, nseg;
  long long nwork;      // padded work index range
  long long unit_start, nunits;  // bra-uniform wavefront units
  long long task_start, ntasks;  // row tasks (workgroups)
  int acc_colsC, acc_colsD;      // LDS row accumulators (0: direct atomics)
  int wpg;                       // wavefronts per workgroup (4 or 8)
  long long nquartets;  // Schwarz-screened quartets
  long long nprim_quartets;
  long long nprim_roots;
  long long nprim_quartets0;  // primitive quartets before primitive-pair screening (EFM-v1's K)
  double cost;          // estimated work (EFM-v1 FLOPs before density screening), or measured ms once tuned
  int stream;           // 0: the caller's stream, k > 0: the context's side stream k - 1
  int task_units;       // wave units per row task (at most)
  std::vector<std::pair<int, int>> rows;  // (first unit relative to unit_start, units) per bra row
  RowTask* d_tasks_own = nullptr;         // row tasks re-cut by the autotuner (owned by the context)
  int task_units0 = 0;                    // the cut stored in the context's d_tasks
  long long ntasks0 = 0;
  std::vector<std::pair<int, std::pair<RowTask*, long long>>> cuts;  // other cuts made so far
  double tuned_ms = 0.0;                  // time of the class pair's launch when tuned (0: not tuned)
  long long rtask_off = 0;                // first of its remapped row tasks (routed mixed builds)
};

// Row tasks of a class pair: each row's wave units cut into
// ceil(nu / task_units) tasks of balanced size, longest (in primitive
// quartets) first, since workgroups pull them from a queue in order and the
// tail of the launch is then made of short tasks.
static std::vector<RowTask> make_tasks(const ClassPairHost& cp, const WaveUnit* units, const std::vector<int>& col,
                                       const std::vector<int>& pp_count, int task_units) {
  std::vector<RowTask> tasks;
  for (const auto& r : cp.rows) {
    const int nu = r.second;
    const int ntk = (nu + task_units - 1) / task_units;
    for (int k = 0; k < ntk; ++k) {
      const int b0 = (int)((long long)nu * k / ntk), b1 = (int)((long long)nu * (k + 1) / ntk);
      tasks.push_back(RowTask{r.first + b0, b1 - b0});
    }
  }
  std::stable_sort(tasks.begin(), tasks.end(), [&](const RowTask& x, const RowTask& y) {
    const WaveUnit& ux = units[x.u0];
    const WaveUnit& uy = units[y.u0];
    const long long cx = (long long)x.nu * pp_count[ux.P] * pp_count[col[ux.col]];
    const long long cy = (long long)y.nu * pp_count[uy.P] * pp_count[col[uy.col]];
    return cx > cy;
  });
  return tasks;
}

__global__ void k_fill(double* x, long long n, double v) {
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x)
    x[i] = v;
}

__global__ void k_zero_u32(unsigned int* x, long long n) {
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x)
    x[i] = 0u;
}

__global__ void k_zero_u64(unsigned long long* x, long long n) {
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x)
    x[i] = 0ull;
}

__global__ void k_copy(double* dst, const double* src, long long n) {
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x)
    dst[i] = src[i];
}

// X <- (X + X^T)/2 in place; thread (i, j) with j > i handles both entries.
__global__ void k_symmetrize(double* X, int n) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  const int i = blockIdx.y * blockDim.y + threadIdx.y;
  if (i >= n || j >= n || j <= i) return;
  const double a = X[(long long)i * n + j], c = X[(long long)j * n + i];
  const double m = 0.5 * (a + c);
  X[(long long)i * n + j] = m;
  X[(long long)j * n + i] = m;
}

// max |D| over each shell-pair block, both orientations (asuka/hf/direct_jk.py:1484-1527);
// with a second density D2 (two-density builds) the max over both.
__global__ void k_dsp(const double* D, const double* D2, DevBasis b, double* Dsp, double* kd,
                      const int* __restrict__ pos_of_sp) {
  const int sp = blockIdx.x * blockDim.x + threadIdx.x;
  if (sp >= b.nsp) return;
  const int A = b.sp_A[sp], B = b.sp_B[sp];
  long long N, msh;
  mol_layout(b, A, N, msh);
  const int a0 = b.shell_ao[A], b0 = b.shell_ao[B];
  const int nA = ncart_c(b.shell_l[A]), nB = ncart_c(b.shell_l[B]);
  const double* X = D + msh;
  const double* X2 = D2 ? D2 + msh : nullptr;
  double m = 0.0;
  for (int i = 0; i < nA; ++i)
    for (int j = 0; j < nB; ++j) {
      m = fmax(m, fabs(X[(a0 + i) * N + b0 + j]));
      if (A != B) m = fmax(m, fabs(X[(b0 + j) * N + a0 + i]));
      if (X2) {
        m = fmax(m, fabs(X2[(a0 + i) * N + b0 + j]));
        if (A != B) m = fmax(m, fabs(X2[(b0 + j) * N + a0 + i]));
      }
    }
  Dsp[sp] = m;
  if (kd && pos_of_sp[sp] >= 0) kd[pos_of_sp[sp]] = m;  // the same value by ket position
}

// Ket-position buffers (EriOut::jblk), one 64-thread workgroup per block:
// Jpos zeroed and Dpos[.. c ..] = D[c-th component of the pair's block] for
// the pair at each position (0 for padding); second density at offset len.
__global__ void k_pos_gather(DevBasis b, const int* __restrict__ col, const long long* __restrict__ jblk,
                             const double* __restrict__ D, const double* __restrict__ D2, double* Dpos,
                             double* Jpos, long long len) {
  const long long k = blockIdx.x;
  const int lane = threadIdx.x;
  const long long base = jblk[k];
  const int ncd = (int)((jblk[k + 1] - base) >> 6);
  const int sp = col[k * 64 + lane];
  long long N = 0, msh = 0;
  int aoC = 0, aoD = 0, nd = 1;
  if (sp >= 0) {
    const int C = b.sp_A[sp], Dh = b.sp_B[sp];
    mol_layout(b, C, N, msh);
    aoC = b.shell_ao[C];
    aoD = b.shell_ao[Dh];
    nd = ncart_c(b.shell_l[Dh]);
  }
  for (int c = 0; c < ncd; ++c) {
    const long long i = base + 64LL * c + lane;
    const long long e = msh + (aoC + c / nd) * N + aoD + c % nd;
    Jpos[i] = 0.0;
    Dpos[i] = sp >= 0 ? D[e] : 0.0;
    if (D2) {
      Jpos[len + i] = 0.0;
      Dpos[len + i] = sp >= 0 ? D2[e] : 0.0;
    }
  }
}

// After the launches: J[c,d] += Jpos (and J2 += the second set). Each element
// of J belongs to one ket position, and no other kernel of the build runs
// any more, so the adds need no atomics.
__global__ void k_pos_scatter(DevBasis b, const int* __restrict__ col, const long long* __restrict__ jblk,
                              const double* __restrict__ Jpos, double* J, double* J2, long long len, int fixed) {
  const long long k = blockIdx.x;
  const int lane = threadIdx.x;
  const int sp = col[k * 64 + lane];
  if (sp < 0) return;
  const long long base = jblk[k];
  const int ncd = (int)((jblk[k + 1] - base) >> 6);
  const int C = b.sp_A[sp], Dh = b.sp_B[sp];
  long long N, msh;
  mol_layout(b, C, N, msh);
  const int aoC = b.shell_ao[C], aoD = b.shell_ao[Dh], nd = ncart_c(b.shell_l[Dh]);
  for (int c = 0; c < ncd; ++c) {
    const long long i = base + 64LL * c + lane;
    const long long e = msh + (aoC + c / nd) * N + aoD + c % nd;
    if (fixed) {  // 64-bit fixed-point sums (EriOut::fixed)
      long long* Jf = reinterpret_cast<long long*>(J);
      long long* J2f = reinterpret_cast<long long*>(J2);
      const long long* Pf = reinterpret_cast<const long long*>(Jpos);
      if (J) Jf[e] = (long long)((unsigned long long)Jf[e] + (unsigned long long)Pf[i]);
      if (J2) J2f[e] = (long long)((unsigned long long)J2f[e] + (unsigned long long)Pf[len + i]);
    } else {
      if (J) J[e] += Jpos[i];
      if (J2) J2[e] += Jpos[len + i];
    }
  }
}

// X[i][j] += Xc[i][g(j)]: the permuted-column flush buffer (EriOut::Kc) into
// the output, g(j) the compact column of AO j.
__global__ void k_unperm_add(double* X, const double* __restrict__ Xc, const int* __restrict__ ao_g, int n,
                             int fixed) {
  const long long nn = (long long)n * n;
  for (long long t = blockIdx.x * (long long)blockDim.x + threadIdx.x; t < nn; t += (long long)gridDim.x * blockDim.x) {
    const long long i = t / n, j = t - i * n;
    if (fixed) {  // 64-bit fixed-point sums (EriOut::fixed)
      unsigned long long* Xf = reinterpret_cast<unsigned long long*>(X);
      Xf[t] += reinterpret_cast<const unsigned long long*>(Xc)[i * n + ao_g[j]];
    } else {
      X[t] += Xc[i * n + ao_g[j]];
    }
  }
}

// Deterministic builds: X (n doubles) to the 64-bit fixed point of the build
// (units of 2^-44, stored in place) and back.
__global__ void k_to_fixed(double* X, long long n) {
  for (long long t = blockIdx.x * (long long)blockDim.x + threadIdx.x; t < n; t += (long long)gridDim.x * blockDim.x)
    X[t] = __longlong_as_double(__double2ll_rn(X[t] * kFixScale));
}
__global__ void k_from_fixed(double* X, long long n) {
  for (long long t = blockIdx.x * (long long)blockDim.x + threadIdx.x; t < n; t += (long long)gridDim.x * blockDim.x)
    X[t] = (double)__double_as_longlong(X[t]) * (1.0 / kFixScale);
}

// out = (X + X^T) / 2 for an n x n block X (batch builds: the packed block of
// one molecule to the caller's matrix).
__global__ void k_sym_copy(double* out, const double* X, int n) {
  const long long nn = (long long)n * n;
  for (long long t = blockIdx.x * (long long)blockDim.x + threadIdx.x; t < nn; t += (long long)gridDim.x * blockDim.x) {
    const long long i = t / n, j = t - i * n;
    out[t] = 0.5 * (X[t] + X[j * n + i]);
  }
}

// Running max of the ket pairs' max|D| within each 64-position block (one
// wavefront per block): dmax[p] = max over positions block_start..p.
__global__ void k_pos_dmax(const int* __restrict__ col, const double* __restrict__ Dsp, double* dmax) {
  const long long p = (long long)blockIdx.x * 64 + threadIdx.x;
  const int sp = col[p];
  double v = sp >= 0 ? Dsp[sp] : 0.0;
#pragma unroll
  for (int off = 1; off < 64; off <<= 1) {
    const double o = __shfl_up(v, off);
    if ((int)threadIdx.x >= off) v = fmax(v, o);
  }
  dmax[p] = v;
}

// Routing of a mixed build, one thread per wave unit, from an upper bound of
// its quartets' bounds Q_ab Q_cd max(D_ab, D_cd): (Q_ab Q_first) max(D_ab,
// max D_cd) with Q_first the unit's largest ket bound (its first position; ket
// lists are Q-descending) and max D_cd the running block max at its last
// position. f32[u] = 1 when that bound is below tau (every quartet of the
// unit is then below tau: the FP32 kernel's share), f64[u] = 1 otherwise;
// both 0 when it is below eps_density (every quartet density-screened). A
// unit kept by the bound whose quartets are all screened is skipped in the
// kernel, which tests each quartet itself.
__global__ void k_route_units(DevBasis b, const WaveUnit* __restrict__ units, const double* __restrict__ qcol,
                              const double* __restrict__ dmaxcol, long long nunits, double eps_density, double tau,
                              int* f64, int* f32) {
  for (long long u = blockIdx.x * (long long)blockDim.x + threadIdx.x; u < nunits;
       u += (long long)gridDim.x * blockDim.x) {
    const WaveUnit un = units[u];
    const double hi = (b.Q[un.P] * qcol[un.col]) * fmax(b.Dsp[un.P], dmaxcol[un.col + un.ncol - 1]);
    const bool any = un.ncol > 0 && (eps_density > 0.0 ? hi >= eps_density : true);
    const bool lowp = hi < tau;
    f32[u] = (any && lowp) ? 1 : 0;
    f64[u] = (any && !lowp) ? 1 : 0;
  }
}

// The routed units, in their original order, into the FP64 and FP32 lists
// (pos: exclusive sums of the flags).
__global__ void k_scatter_units(const WaveUnit* __restrict__ units, long long nunits, const int* __restrict__ f64,
                                const int* __restrict__ f32, const int* __restrict__ pos64,
                                const int* __restrict__ pos32, WaveUnit* u64, WaveUnit* u32) {
  for (long long u = blockIdx.x * (long long)blockDim.x + threadIdx.x; u < nunits; u += (long long)gridDim.x * blockDim.x) {
    if (f32[u]) u32[pos32[u]] = units[u];
    if (f64[u]) u64[pos64[u]] = units[u];
  }
}

// A class pair's row tasks (units [ubase + u0, ubase + u0 + nu) of the full
// list) as the ranges their units occupy in the two routed lists; a range may
// be empty.
__global__ void k_route_tasks(const RowTask* __restrict__ tasks, long long ntasks, long long ubase,
                              const int* __restrict__ pos64, const int* __restrict__ pos32, RowTask* t64,
                              RowTask* t32) {
  for (long long t = blockIdx.x * (long long)blockDim.x + threadIdx.x; t < ntasks; t += (long long)gridDim.x * blockDim.x) {
    const RowTask tk = tasks[t];
    const long long g0 = ubase + tk.u0, g1 = g0 + tk.nu;
    t64[t] = RowTask{pos64[g0], pos64[g1] - pos64[g0]};
    t32[t] = RowTask{pos32[g0], pos32[g1] - pos32[g0]};
  }
}

__global__ void k_rys_points(int n, long long npts, const double* T, double* x, double* w) {
  const long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  if (i >= npts) return;
  for (int u = 0; u < n; ++u) rys_root_dev(n, u, T[i], x[i * n + u], w[i * n + u]);
}

template <typename Tp>
int dev_upload(Tp** dst, const std::vector<Tp>& src, std::vector<void*>& owned, hipStream_t s) {
  *dst = nullptr;
  size_t bytes = sizeof(Tp) * (src.empty() ? 1 : src.size());
  hipError_t e = hipMalloc((void**)dst, bytes);
  if (e != hipSuccess) return fail(ASUKA_E_NOMEM, std::string("hipMalloc: ") + hipGetErrorString(e));
  owned.push_back(*dst);
  if (!src.empty()) {
    // blocking copy: callers pass temporaries (context construction is one-time)
    (void)s;
    e = hipMemcpy(*dst, src.data(), sizeof(Tp) * src.size(), hipMemcpyHostToDevice);
    if (e != hipSuccess) return fail(ASUKA_E_HIP, std::string("hipMemcpy: ") + hipGetErrorString(e));
  }
  return 0;
}

}  // namespace

hipError_t launch_int1e(const asuka_jk_ctx* ctx, double* S, double* T, double* V, hipStream_t s);
hipError_t launch_eri_deriv(const DevBasis& b, const ClassWork& w, const double* D, const double* WJ,
                            const double* WK, const int* shell_atom, double* grad, hipStream_t s);
hipError_t launch_eri_deriv_fast(const DevBasis& b, const ClassWork& w, const double* D, const double* WJ,
                                 const double* WK, const int* shell_atom, int natm, double* grad, hipStream_t s,
                                 bool* handled);
hipError_t launch_int1e_grad(const DevBasis& b, const int* pstart, const int* pnum, const double* pexp,
                             const double* pcoef, int natm, const double* axyz, const double* aZ,
                             const int* shell_atom, const double* Dh, const double* Ws, double* grad,
                             hipStream_t s);

struct asuka_jk_ctx {
  int device = 0;
  int nshell = 0, nao = 0, nsp = 0, natm = 0, max_l = 0;
  double eps_schwarz = 0, eps_density = 0, tau_mixed = 0;
  std::vector<int> shell_l, shell_ao, shell_atom, sp_A, sp_B, pp_start, pp_count;
  std::vector<int> pp_count0;  // primitive pairs per shell pair before primitive-pair screening
  std::vector<int> shell_mol;  // molecule of each shell (batch contexts; all 0 otherwise)
  int nmol = 1;
  std::vector<int> mol_ao_off, mol_nao;  // each molecule's first AO and AO count
  long long nelem = 0;                   // elements of a D/J/K buffer (sum of the molecules' n^2)
  double *d_Dbig = nullptr, *d_Jbig = nullptr, *d_Kbig = nullptr;  // batch contexts: block-diagonal D, J, K
  std::vector<double> shell_xyz, Q, atom_xyz, atom_Z;
  std::vector<PrimPair> pp;
  std::vector<PrimPairF> ppf;
  PrimPairF32* d_ppf32 = nullptr;  // float copy of the final ppf (FP32 path)
  std::vector<ClassPairHost> cps;
  long long ntasks = 0;
  long long nprim_dropped = 0;  // primitive pairs removed by primitive-pair screening
  DevBasis dev{};
  int* d_row_pair = nullptr;
  long long* d_row_off = nullptr;
  int* d_col_pair = nullptr;
  Segment* d_segs = nullptr;
  WaveUnit* d_units = nullptr;
  RowTask* d_tasks = nullptr;
  std::vector<WaveUnit> h_units;  // host copies of the wave units and ket lists
  // routing of mixed builds: per wave unit FP64/FP32 flags (nunits + 1
  // entries, the last 0), their exclusive sums, the two routed unit lists, the
  // remapped row tasks of every class pair (rtask_cap entries each) and the
  // scan scratch
  int *d_rf64 = nullptr, *d_rf32 = nullptr, *d_rp64 = nullptr, *d_rp32 = nullptr;
  // per ket position: Q of the pair (0 for padding) and, per build, the
  // running max of the pairs' max|D| from the position's 64-block start
  double *d_qcol = nullptr, *d_dmaxcol = nullptr;
  double* d_kd = nullptr;          // DevBasis::kd
  int* d_pos_of_sp = nullptr;      // position of each shell pair in the ket lists (-1: none)
  std::vector<int> h_shell_lao;    // host copy of DevBasis::shell_lao
  WaveUnit *d_units64 = nullptr, *d_units32 = nullptr;
  RowTask *d_rtasks64 = nullptr, *d_rtasks32 = nullptr;
  long long rtask_cap = 0;
  void* d_scan_tmp = nullptr;
  size_t scan_tmp_bytes = 0;
  // ket-position buffers (EriOut::jblk): nblk + 1 block offsets, then J[c,d]
  // and D[c,d] for two densities (2 * pos_len doubles each)
  long long nblk = 0, pos_len = 0;
  long long* d_jblk = nullptr;
  double *d_jpos = nullptr, *d_dpos = nullptr;
  // permuted-column K of the row-task flush (EriOut::Kc; one-molecule
  // contexts): two nao x nao buffers and the compact column of each AO
  double* d_kc = nullptr;
  int* d_ao_g = nullptr;
  std::vector<int> h_col;
  int lao_off[ASUKA_LMAX + 2] = {0};
  size_t lds_per_group = 64 * 1024;  // device limit of LDS per workgroup
  size_t row_acc_max = kRowAccMaxBytes;  // LDS row-accumulator budget (bytes)
  double eps_prim = kPrimQuartetEps;     // primitive-quartet cut of the register kernels
  int nlao[ASUKA_LMAX + 1] = {0};
  double* d_Q = nullptr;
  double* d_Dsp = nullptr;
  double* d_atom_xyz = nullptr;
  double* d_atom_Z = nullptr;
  int* d_shell_atom = nullptr;
  int* d_prim_start = nullptr;
  int* d_prim_num = nullptr;
  double* d_prim_exp = nullptr;
  double* d_prim_coef = nullptr;
  unsigned long long* d_counters = nullptr;
  unsigned int* d_queue = nullptr;  // one work-queue head per class pair
  hipEvent_t ev_start = nullptr, ev_stop = nullptr;
  // Side streams (default priority) over which the class-pair launches of a
  // build are spread, forked from and joined back into the caller's stream
  // by events, so one launch's tail overlaps the next launches.
  std::vector<hipStream_t> side;
  hipEvent_t ev_fork = nullptr;
  std::vector<hipEvent_t> ev_join;
  std::vector<int> launch_order;  // class pairs, largest estimated cost first
  long long last_launches = 0;
  std::vector<void*> owned;
  ~asuka_jk_ctx() {
    if (ev_start) (void)hipEventDestroy(ev_start);
    if (ev_stop) (void)hipEventDestroy(ev_stop);
    if (ev_fork) (void)hipEventDestroy(ev_fork);
    for (hipEvent_t e : ev_join) (void)hipEventDestroy(e);
    for (hipStream_t st : side) (void)hipStreamDestroy(st);
    for (void* p : owned) (void)hipFree(p);
  }
};

// Class pairs over streams, largest first, each to the least loaded stream
// (longest-processing-time rule). The weights are the measured launch times
// once every class pair is tuned, else the EFM-v1 estimates.
static void assign_streams(asuka_jk_ctx* ctx) {
  const int nst = 1 + (int)ctx->side.size();
  bool tuned = !ctx->cps.empty();
  for (const auto& cp : ctx->cps) tuned = tuned && cp.tuned_ms > 0.0;
  auto weight = [&](int i) { return tuned ? ctx->cps[i].tuned_ms : ctx->cps[i].cost; };
  std::vector<int> order(ctx->cps.size());
  std::iota(order.begin(), order.end(), 0);
  std::stable_sort(order.begin(), order.end(), [&](int x, int y) { return weight(x) > weight(y); });
  std::vector<double> load(nst, 0.0);
  for (int i : order) {
    const int k = (int)(std::min_element(load.begin(), load.end()) - load.begin());
    ctx->cps[i].stream = k;
    load[k] += weight(i);
  }
  ctx->launch_order = order;
}

// LDS row-accumulator columns of a class pair when they fit (else 0, 0).
static bool row_acc_fits(const asuka_jk_ctx* ctx, int la, int lb, int lc, int ld, int* cC, int* cD) {
  *cC = ctx->nlao[lc];
  *cD = (lc == ld) ? 0 : ctx->nlao[ld];
  const size_t bytes =
      sizeof(double) * (size_t)(ncart_c(la) * ncart_c(lb) + (ncart_c(la) + ncart_c(lb)) * (*cC + *cD));
  const int nr = (la + lb + lc + ld) / 2 + 1;
  return bytes <= std::min(ctx->row_acc_max, ctx->lds_per_group - root_table_lds(nr) - 1024);
}

// Offsets of each class pair's remapped row tasks (of its current cut) in the
// routed task arrays, which grow when a cut needs more room (context creation
// and the autotuner's re-cuts; never during a build).
static int route_task_layout(asuka_jk_ctx* ctx) {
  long long tot = 0;
  for (auto& cp : ctx->cps) {
    cp.rtask_off = tot;
    tot += cp.ntasks;
  }
  tot = std::max<long long>(1, tot);
  if (tot > ctx->rtask_cap) {
    RowTask *a = nullptr, *b = nullptr;
    HIP_CHECK(hipMalloc(&a, sizeof(RowTask) * (size_t)tot));
    ctx->owned.push_back(a);
    HIP_CHECK(hipMalloc(&b, sizeof(RowTask) * (size_t)tot));
    ctx->owned.push_back(b);
    ctx->d_rtasks64 = a;
    ctx->d_rtasks32 = b;
    ctx->rtask_cap = tot;
  }
  return ASUKA_OK;
}

// Device buffers of the routing of mixed builds (k_route_units and after).
// Block offsets of the ket-position buffers (EriOut::jblk) for the padded ket
// lists, and the buffers themselves.
static int pos_layout(asuka_jk_ctx* ctx, const std::vector<int>& col) {
  const long long nblk = (long long)col.size() / 64;
  std::vector<long long> jblk((size_t)nblk + 1, 0);
  for (long long k = 0; k < nblk; ++k) {
    int ncd = 0;
    for (int t = 0; t < 64 && !ncd; ++t) {
      const int sp = col[(size_t)(k * 64 + t)];
      if (sp >= 0) ncd = ncart_c(ctx->shell_l[ctx->sp_A[sp]]) * ncart_c(ctx->shell_l[ctx->sp_B[sp]]);
    }
    jblk[(size_t)k + 1] = jblk[(size_t)k] + 64LL * ncd;
  }
  ctx->nblk = nblk;
  ctx->pos_len = jblk.back();
  int rc = dev_upload(&ctx->d_jblk, jblk, ctx->owned, nullptr);
  if (rc) return rc;
  std::vector<double> qcol(col.size(), 0.0), kcd(3 * col.size(), 0.0);
  std::vector<int2> kpp(col.size(), int2{0, 0}), kao(col.size(), int2{0, 0}), klao(col.size(), int2{0, 0});
  std::vector<int> pos_of_sp(ctx->nsp, -1);
  for (size_t t = 0; t < col.size(); ++t) {
    const int sp = col[t];
    if (sp < 0) continue;
    const int C = ctx->sp_A[sp], Dh = ctx->sp_B[sp];
    qcol[t] = ctx->Q[sp];
    kpp[t] = int2{ctx->pp_start[sp], ctx->pp_count[sp]};
    kao[t] = int2{ctx->shell_ao[C], ctx->shell_ao[Dh]};
    klao[t] = int2{ctx->h_shell_lao[C], ctx->h_shell_lao[Dh]};
    for (int x = 0; x < 3; ++x) kcd[3 * t + x] = ctx->shell_xyz[3 * C + x] - ctx->shell_xyz[3 * Dh + x];
    pos_of_sp[sp] = (int)t;
  }
  if ((rc = dev_upload(&ctx->d_qcol, qcol, ctx->owned, nullptr))) return rc;
  int2 *d_kpp, *d_kao, *d_klao;
  double* d_kcd;
  if ((rc = dev_upload(&d_kpp, kpp, ctx->owned, nullptr))) return rc;
  if ((rc = dev_upload(&d_kao, kao, ctx->owned, nullptr))) return rc;
  if ((rc = dev_upload(&d_klao, klao, ctx->owned, nullptr))) return rc;
  if ((rc = dev_upload(&d_kcd, kcd, ctx->owned, nullptr))) return rc;
  if ((rc = dev_upload(&ctx->d_pos_of_sp, pos_of_sp, ctx->owned, nullptr))) return rc;
  std::vector<double> kd0(col.size(), 0.0);  // padding positions stay 0
  if ((rc = dev_upload(&ctx->d_kd, kd0, ctx->owned, nullptr))) return rc;
  {  // transposed ket primitive pairs (DevBasis::kpoff)
    const long long nb = (long long)col.size() / 64;
    std::vector<long long> kpoff((size_t)nb + 1, 0);
    for (long long k = 0; k < nb; ++k) {
      int kmax = 0;
      for (int t = 0; t < 64; ++t) {
        const int sp = col[(size_t)(k * 64 + t)];
        if (sp >= 0) kmax = std::max(kmax, ctx->pp_count[sp]);
      }
      kpoff[(size_t)k + 1] = kpoff[(size_t)k] + kmax;
    }
    const long long nslot = std::max<long long>(1, kpoff.back() * 64);
    if (nslot >= (1LL << 31)) return fail(ASUKA_E_INVALID, "ket primitive-pair table exceeds 2^31 slots");
    std::vector<double> k64((size_t)(11 * nslot), 0.0), k32d((size_t)(3 * nslot), 0.0);
    std::vector<float> k32f((size_t)(7 * nslot), 0.0f);
    for (long long k = 0; k < nb; ++k)
      for (int t = 0; t < 64; ++t) {
        const int sp = col[(size_t)(k * 64 + t)];
        if (sp < 0) continue;
        for (int j = 0; j < ctx->pp_count[sp]; ++j) {
          const PrimPairF& f = ctx->ppf[(size_t)ctx->pp_start[sp] + j];
          const double a_exp = ctx->pp[(size_t)ctx->pp_start[sp] + j].a_exp;
          const size_t slot = (size_t)((kpoff[(size_t)k] + j) * 64 + t), n = (size_t)nslot;
          const double v64[11] = {f.eta, f.inv_eta, f.ck, f.Px, f.Py, f.Pz, f.PAx, f.PAy, f.PAz, f.q, a_exp};
          for (int x = 0; x < 11; ++x) k64[x * n + slot] = v64[x];
          k32d[0 * n + slot] = f.Px;
          k32d[1 * n + slot] = f.Py;
          k32d[2 * n + slot] = f.Pz;
          // the same roundings as the FP32 pair table (q clamped to the float range)
          const float v32[7] = {(float)f.eta, (float)f.inv_eta, (float)f.ck, (float)f.PAx, (float)f.PAy, (float)f.PAz,
                                (float)std::min(f.q, 3e38)};
          for (int x = 0; x < 7; ++x) k32f[x * n + slot] = v32[x];
        }
      }
    long long* d_kpoff;
    double *d_k64, *d_k32d;
    float* d_k32f;
    if ((rc = dev_upload(&d_kpoff, kpoff, ctx->owned, nullptr))) return rc;
    if ((rc = dev_upload(&d_k64, k64, ctx->owned, nullptr))) return rc;
    if ((rc = dev_upload(&d_k32d, k32d, ctx->owned, nullptr))) return rc;
    if ((rc = dev_upload(&d_k32f, k32f, ctx->owned, nullptr))) return rc;
    ctx->dev.kpoff = d_kpoff;
    ctx->dev.kp64 = d_k64;
    ctx->dev.kp32d = d_k32d;
    ctx->dev.kp32f = d_k32f;
    ctx->dev.kslots = nslot;
  }
  ctx->dev.kq = ctx->d_qcol;
  ctx->dev.kd = ctx->d_kd;
  ctx->dev.kpp = d_kpp;
  ctx->dev.kao = d_kao;
  ctx->dev.klao = d_klao;
  ctx->dev.kcd = d_kcd;
  HIP_CHECK(hipMalloc((void**)&ctx->d_dmaxcol, sizeof(double) * std::max<size_t>(1, col.size())));
  ctx->owned.push_back(ctx->d_dmaxcol);
  const size_t bytes = sizeof(double) * (size_t)std::max<long long>(1, 2 * ctx->pos_len);
  HIP_CHECK(hipMalloc((void**)&ctx->d_jpos, bytes));
  ctx->owned.push_back(ctx->d_jpos);
  HIP_CHECK(hipMalloc((void**)&ctx->d_dpos, bytes));
  ctx->owned.push_back(ctx->d_dpos);
  return ASUKA_OK;
}

static int route_alloc(asuka_jk_ctx* ctx) {
  const long long n = (long long)ctx->h_units.size();
  const size_t ni = sizeof(int) * (size_t)(n + 1);
  int** flags[4] = {&ctx->d_rf64, &ctx->d_rf32, &ctx->d_rp64, &ctx->d_rp32};
  for (int** f : flags) {
    HIP_CHECK(hipMalloc((void**)f, ni));
    ctx->owned.push_back(*f);
    HIP_CHECK(hipMemset(*f, 0, ni));  // context creation, not a build: the flags' last entry stays 0
  }
  const size_t nu = sizeof(WaveUnit) * (size_t)std::max<long long>(1, n);
  HIP_CHECK(hipMalloc((void**)&ctx->d_units64, nu));
  ctx->owned.push_back(ctx->d_units64);
  HIP_CHECK(hipMalloc((void**)&ctx->d_units32, nu));
  ctx->owned.push_back(ctx->d_units32);
  HIP_CHECK(hipcub::DeviceScan::ExclusiveSum(nullptr, ctx->scan_tmp_bytes, ctx->d_rf64, ctx->d_rp64, (int)(n + 1)));
  ctx->scan_tmp_bytes = std::max<size_t>(ctx->scan_tmp_bytes, 256);
  HIP_CHECK(hipMalloc(&ctx->d_scan_tmp, ctx->scan_tmp_bytes));
  ctx->owned.push_back(ctx->d_scan_tmp);
  return route_task_layout(ctx);
}

extern "C" {

int asuka_jk_version(void) { return ASUKA_JK_ABI_VERSION; }

const char* asuka_jk_last_error(void) { return g_err.c_str(); }

}  // extern "C"

// Builds a context for the shells of bd. shell_mol (optional) assigns each
// shell to one of several independent molecules: shell pairs, quartets and
// the LDS accumulator columns then stay within a molecule (a batch context).
static int ctx_build(const asuka_basis_desc* bd, const std::vector<int>* shell_mol, const asuka_jk_opts* opts,
                     hipStream_t s, asuka_jk_ctx** out) {
  if (!bd || !opts || !out) return fail(ASUKA_E_INVALID, "basis, opts and out must be non-null");
  *out = nullptr;
  const int nshell = bd->nshell;
  if (nshell <= 0) return fail(ASUKA_E_INVALID, "AO basis has zero shells");
  if (!bd->shell_cxyz || !bd->shell_prim_start || !bd->shell_nprim || !bd->shell_l || !bd->shell_ao_start ||
      !bd->prim_exp || !bd->prim_coef)
    return fail(ASUKA_E_INVALID, "basis arrays must be non-null");
  if (!(opts->eps_schwarz >= 0.0)) return fail(ASUKA_E_INVALID, "eps_schwarz must be >= 0");

  auto* ctx = new asuka_jk_ctx();
  // on a failed build the context is freed only after the work it queued has
  // finished (as asuka_jk_ctx_destroy does)
  struct SyncDelete {
    void operator()(asuka_jk_ctx* c) const {
      (void)hipDeviceSynchronize();
      delete c;
    }
  };
  std::unique_ptr<asuka_jk_ctx, SyncDelete> guard(ctx);
  HIP_CHECK(hipGetDevice(&ctx->device));
  {
    int v = 0;
    if (hipDeviceGetAttribute(&v, hipDeviceAttributeMaxSharedMemoryPerBlock, ctx->device) == hipSuccess && v > 0)
      ctx->lds_per_group = (size_t)v;
    if (const char* e = std::getenv("ASUKA_EPS_PRIM"); e && *e) {  // measurement of the cut's cost
      const double v = std::strtod(e, nullptr);
      if (v >= 0.0 && v <= 1e-10) ctx->eps_prim = v;
    }
    if (const char* e = std::getenv("ASUKA_ROW_ACC_MAX_KB"); e && *e) {
      const long kb = std::strtol(e, nullptr, 10);
      if (kb >= 0) ctx->row_acc_max = std::min(kRowAccMaxBytes, (size_t)kb * 1024);
    }
  }
  ctx->nshell = nshell;
  ctx->shell_l.assign(bd->shell_l, bd->shell_l + nshell);
  ctx->shell_ao.assign(bd->shell_ao_start, bd->shell_ao_start + nshell);
  ctx->shell_xyz.assign(bd->shell_cxyz, bd->shell_cxyz + 3 * nshell);
  for (int i = 0; i < nshell; ++i) {
    const int l = ctx->shell_l[i];
    if (l < 0) return fail(ASUKA_E_INVALID, "shell_l must be >= 0");
    if (l > ASUKA_LMAX)
      return fail(ASUKA_E_UNSUPPORTED, "direct J/K supports only l <= " + std::to_string(ASUKA_LMAX));
    ctx->max_l = std::max(ctx->max_l, l);
    const int ps = bd->shell_prim_start[i], np = bd->shell_nprim[i];
    if (np <= 0 || ps < 0 || ps + np > bd->nprim) return fail(ASUKA_E_INVALID, "shell primitive range out of bounds");
  }
  ctx->nao = ctx->shell_ao[nshell - 1] + ncart_c(ctx->shell_l[nshell - 1]);
  if (shell_mol) {
    if ((int)shell_mol->size() != nshell) return fail(ASUKA_E_INVALID, "shell_mol must have nshell entries");
    ctx->shell_mol = *shell_mol;
  } else {
    ctx->shell_mol.assign(nshell, 0);
  }
  ctx->nmol = 1 + *std::max_element(ctx->shell_mol.begin(), ctx->shell_mol.end());
  const int nmol = ctx->nmol;
  ctx->natm = bd->natm;
  if (bd->natm > 0) {
    if (!bd->atom_xyz || !bd->atom_charge) return fail(ASUKA_E_INVALID, "atom arrays must be non-null when natm > 0");
    ctx->atom_xyz.assign(bd->atom_xyz, bd->atom_xyz + 3 * bd->natm);
    ctx->atom_Z.assign(bd->atom_charge, bd->atom_charge + bd->natm);
  }
  if (bd->shell_atom)
    ctx->shell_atom.assign(bd->shell_atom, bd->shell_atom + nshell);
  else
    ctx->shell_atom.assign(nshell, 0);
  ctx->eps_schwarz = opts->eps_schwarz;
  ctx->eps_density = opts->eps_density < 0.0 ? std::max(opts->eps_schwarz, 1e-10) : opts->eps_density;
  ctx->tau_mixed = opts->tau_mixed > 0.0 ? opts->tau_mixed : kDefaultTauMixed;
  {
    const int nst = opts->streams > 0 ? std::min(opts->streams, kMaxStreams) : kDefaultStreams;
    HIP_CHECK(hipEventCreateWithFlags(&ctx->ev_fork, hipEventDisableTiming));
    for (int k = 1; k < nst; ++k) {
      hipStream_t st = nullptr;
      HIP_CHECK(hipStreamCreateWithFlags(&st, hipStreamNonBlocking));
      ctx->side.push_back(st);
      hipEvent_t e = nullptr;
      HIP_CHECK(hipEventCreateWithFlags(&e, hipEventDisableTiming));
      ctx->ev_join.push_back(e);
    }
  }

  // ---- shell pairs, l-ordered orientation (asuka/cueri/shell_pairs.py:72-116) ----
  for (int i = 0; i < nshell; ++i)
    for (int j = 0; j <= i; ++j) {
      if (ctx->shell_mol[i] != ctx->shell_mol[j]) continue;  // no pairs across the molecules of a batch
      int A = i, B = j;
      if (ctx->shell_l[i] < ctx->shell_l[j]) std::swap(A, B);
      ctx->sp_A.push_back(A);
      ctx->sp_B.push_back(B);
    }
  ctx->nsp = (int)ctx->sp_A.size();

  // ---- primitive pairs (pyx:2066-2129), exact zeros dropped ----
  for (int sp = 0; sp < ctx->nsp; ++sp) {
    const int A = ctx->sp_A[sp], B = ctx->sp_B[sp];
    const double* a3 = &ctx->shell_xyz[3 * A];
    const double* b3 = &ctx->shell_xyz[3 * B];
    const double dx = a3[0] - b3[0], dy = a3[1] - b3[1], dz = a3[2] - b3[2];
    const double AB2 = dx * dx + dy * dy + dz * dz;
    ctx->pp_start.push_back((int)ctx->pp.size());
    for (int ia = 0; ia < bd->shell_nprim[A]; ++ia)
      for (int ib = 0; ib < bd->shell_nprim[B]; ++ib) {
        const int pA = bd->shell_prim_start[A] + ia, pB = bd->shell_prim_start[B] + ib;
        const double al = bd->prim_exp[pA], be = bd->prim_exp[pB];
        const double eta = al + be, ie = 1.0 / eta;
        PrimPair r;
        r.eta = eta;
        r.Px = (al * a3[0] + be * b3[0]) * ie;
        r.Py = (al * a3[1] + be * b3[1]) * ie;
        r.Pz = (al * a3[2] + be * b3[2]) * ie;
        r.cK = bd->prim_coef[pA] * bd->prim_coef[pB] * std::exp(-(al * be) * ie * AB2);
        r.a_exp = al;
        r.b_exp = be;
        r.inv_eta = ie;
        if (r.cK != 0.0) {
          ctx->pp.push_back(r);
          PrimPairF f{};
          f.eta = eta;
          f.inv_eta = ie;
          f.ck = kSqrt2Pi52 * r.cK * ie;
          f.Px = r.Px;
          f.Py = r.Py;
          f.Pz = r.Pz;
          f.PAx = r.Px - a3[0];
          f.PAy = r.Py - a3[1];
          f.PAz = r.Pz - a3[2];
          f.q = 1e300;  // no bound yet (set by primitive-pair screening)
          ctx->ppf.push_back(f);
        }
      }
    ctx->pp_count.push_back((int)ctx->pp.size() - ctx->pp_start.back());
  }

  // ---- device copies ----
  std::vector<void*>& own = ctx->owned;
  int rc;
  double* d_xyz;
  int *d_l, *d_ao, *d_spA, *d_spB, *d_pps, *d_ppc;
  PrimPair* d_pp;
  PrimPairF* d_ppf;
  if ((rc = dev_upload(&d_xyz, ctx->shell_xyz, own, s))) return rc;
  if ((rc = dev_upload(&d_l, ctx->shell_l, own, s))) return rc;
  if ((rc = dev_upload(&d_ao, ctx->shell_ao, own, s))) return rc;
  if ((rc = dev_upload(&d_spA, ctx->sp_A, own, s))) return rc;
  if ((rc = dev_upload(&d_spB, ctx->sp_B, own, s))) return rc;
  if ((rc = dev_upload(&d_pps, ctx->pp_start, own, s))) return rc;
  if ((rc = dev_upload(&d_ppc, ctx->pp_count, own, s))) return rc;
  if ((rc = dev_upload(&d_pp, ctx->pp, own, s))) return rc;
  if ((rc = dev_upload(&d_ppf, ctx->ppf, own, s))) return rc;
  if ((rc = dev_upload(&ctx->d_shell_atom, ctx->shell_atom, own, s))) return rc;
  {
    std::vector<int> ps(bd->shell_prim_start, bd->shell_prim_start + nshell), pn(bd->shell_nprim, bd->shell_nprim + nshell);
    std::vector<double> pe(bd->prim_exp, bd->prim_exp + bd->nprim), pc(bd->prim_coef, bd->prim_coef + bd->nprim);
    if ((rc = dev_upload(&ctx->d_prim_start, ps, own, s))) return rc;
    if ((rc = dev_upload(&ctx->d_prim_num, pn, own, s))) return rc;
    if ((rc = dev_upload(&ctx->d_prim_exp, pe, own, s))) return rc;
    if ((rc = dev_upload(&ctx->d_prim_coef, pc, own, s))) return rc;
  }
  // compact AO numbering per angular momentum and molecule (the LDS
  // accumulator columns): lao_ao lists, for each l, the AOs of angular
  // momentum l molecule by molecule; shell_lao is a shell's first index in
  // its molecule's list and mol_lao where that list starts within l's part.
  {
    std::vector<int> shell_lao(nshell), lao_ao, mol_lao((size_t)nmol * (ASUKA_LMAX + 1));
    for (int l = 0; l <= ASUKA_LMAX; ++l) {
      ctx->lao_off[l] = (int)lao_ao.size();
      ctx->nlao[l] = 0;
      for (int m = 0; m < nmol; ++m) {
        const int m0 = (int)lao_ao.size();
        mol_lao[(size_t)m * (ASUKA_LMAX + 1) + l] = m0 - ctx->lao_off[l];
        for (int sh = 0; sh < nshell; ++sh)
          if (ctx->shell_l[sh] == l && ctx->shell_mol[sh] == m) {
            shell_lao[sh] = (int)lao_ao.size() - m0;
            for (int c = 0; c < ncart_c(l); ++c) lao_ao.push_back(ctx->shell_ao[sh] + c);
          }
        ctx->nlao[l] = std::max(ctx->nlao[l], (int)lao_ao.size() - m0);  // columns of the largest molecule
      }
    }
    ctx->lao_off[ASUKA_LMAX + 1] = (int)lao_ao.size();
    int *d_slao, *d_laoao, *d_mlao, *d_smol;
    if ((rc = dev_upload(&d_slao, shell_lao, own, s))) return rc;
    if ((rc = dev_upload(&d_laoao, lao_ao, own, s))) return rc;
    if ((rc = dev_upload(&d_mlao, mol_lao, own, s))) return rc;
    if ((rc = dev_upload(&d_smol, ctx->shell_mol, own, s))) return rc;
    ctx->dev.shell_lao = d_slao;
    ctx->h_shell_lao = shell_lao;
    ctx->dev.lao_ao = d_laoao;
    ctx->dev.mol_lao = d_mlao;
    ctx->dev.shell_mol = d_smol;
    if (nmol == 1 && (int)lao_ao.size() == ctx->nao) {  // permuted-column K of the flush (EriOut::Kc)
      std::vector<int> ao_g(ctx->nao);
      for (int g = 0; g < ctx->nao; ++g) ao_g[lao_ao[g]] = g;
      if ((rc = dev_upload(&ctx->d_ao_g, ao_g, own, s))) return rc;
      HIP_CHECK(hipMalloc((void**)&ctx->d_kc, sizeof(double) * 2 * (size_t)ctx->nao * ctx->nao));
      own.push_back(ctx->d_kc);
    }
  }
  // per-molecule D/J/K blocks (a molecule's AOs are contiguous): row-major
  // n_m x n_m blocks packed one after another; one molecule: plain nao x nao
  {
    std::vector<int> lo(nmol, ctx->nao), hi(nmol, 0), cnt(nmol, 0);
    for (int sh = 0; sh < nshell; ++sh) {
      const int m = ctx->shell_mol[sh], n = ncart_c(ctx->shell_l[sh]);
      lo[m] = std::min(lo[m], ctx->shell_ao[sh]);
      hi[m] = std::max(hi[m], ctx->shell_ao[sh] + n);
      cnt[m] += n;
    }
    std::vector<int> mn(nmol);
    std::vector<long long> msh(nmol);
    long long base = 0;
    for (int m = 0; m < nmol; ++m) {
      if (cnt[m] == 0 || hi[m] - lo[m] != cnt[m]) return fail(ASUKA_E_INVALID, "each molecule's AOs must be contiguous");
      mn[m] = cnt[m];
      msh[m] = base - (long long)lo[m] * (cnt[m] + 1);
      base += (long long)cnt[m] * cnt[m];
    }
    ctx->mol_ao_off = lo;
    ctx->mol_nao = mn;
    ctx->nelem = base;
    int* d_mn;
    long long* d_msh;
    if ((rc = dev_upload(&d_mn, mn, own, s))) return rc;
    if ((rc = dev_upload(&d_msh, msh, own, s))) return rc;
    ctx->dev.mol_n = d_mn;
    ctx->dev.mol_shift = d_msh;
  }
  std::vector<double> zeros(ctx->nsp, 0.0);
  if ((rc = dev_upload(&ctx->d_Q, zeros, own, s))) return rc;
  if ((rc = dev_upload(&ctx->d_Dsp, zeros, own, s))) return rc;
  if (ctx->natm > 0) {
    if ((rc = dev_upload(&ctx->d_atom_xyz, ctx->atom_xyz, own, s))) return rc;
    if ((rc = dev_upload(&ctx->d_atom_Z, ctx->atom_Z, own, s))) return rc;
  }
  DevBasis& dv = ctx->dev;
  dv.shell_xyz = d_xyz;
  dv.shell_l = d_l;
  dv.shell_ao = d_ao;
  dv.sp_A = d_spA;
  dv.sp_B = d_spB;
  dv.pp_start = d_pps;
  dv.pp_count = d_ppc;
  dv.pp = d_pp;
  dv.ppf = d_ppf;
  dv.Q = ctx->d_Q;
  dv.Dsp = ctx->d_Dsp;
  dv.nao = ctx->nao;
  dv.nsp = ctx->nsp;

  // ---- primitive-pair screening (only with Schwarz screening on) ----
  // The Schwarz bound of every primitive pair alone, sqrt(max_i (p_i|p_i))
  // including its coefficient product: one diagonal quartet per primitive
  // pair, the generic kernel in Schwarz mode over a table in which each
  // primitive pair is its own shell pair. A primitive pair whose bound times
  // the largest bound is below kPrimPairEps adds less than that to every
  // integral it enters (Cauchy-Schwarz per primitive quartet) and is dropped;
  // these are mostly tight primitives on distant centres, whose
  // exp(-a b / (a + b) |AB|^2) is tiny but not zero.
  if (ctx->eps_schwarz > 0.0 && !ctx->pp.empty()) {
    const int npp = (int)ctx->pp.size();
    std::vector<void*> tmp;
    struct FreeAll {
      std::vector<void*>& v;
      ~FreeAll() {
        for (void* p : v) (void)hipFree(p);
      }
    } free_tmp{tmp};
    std::vector<int> tA(npp), tB(npp), tS(npp), tC(npp, 1);
    for (int sp = 0; sp < ctx->nsp; ++sp)
      for (int t = ctx->pp_start[sp]; t < ctx->pp_start[sp] + ctx->pp_count[sp]; ++t) {
        tA[t] = ctx->sp_A[sp];
        tB[t] = ctx->sp_B[sp];
        tS[t] = t;
      }
    int *d_tA, *d_tB, *d_tS, *d_tC;
    double* d_Qp;
    if ((rc = dev_upload(&d_tA, tA, tmp, s))) return rc;
    if ((rc = dev_upload(&d_tB, tB, tmp, s))) return rc;
    if ((rc = dev_upload(&d_tS, tS, tmp, s))) return rc;
    if ((rc = dev_upload(&d_tC, tC, tmp, s))) return rc;
    if ((rc = dev_upload(&d_Qp, std::vector<double>(npp, 0.0), tmp, s))) return rc;
    DevBasis tb = dv;
    tb.sp_A = d_tA;
    tb.sp_B = d_tB;
    tb.pp_start = d_tS;
    tb.pp_count = d_tC;
    tb.nsp = npp;
    ClassWork w{};
    w.la = w.lb = w.lc = w.ld = ctx->max_l;
    w.nwork = npp;
    w.q_ab = d_tS;  // the identity list: quartet t is (t|t)
    w.q_cd = d_tS;
    EriOut o{};
    o.mode = MODE_SCHWARZ;
    o.Qout = d_Qp;
    HIP_CHECK(launch_eri_generic(tb, w, o, s));
    std::vector<double> Qp(npp);
    HIP_CHECK(hipMemcpyAsync(Qp.data(), d_Qp, sizeof(double) * npp, hipMemcpyDeviceToHost, s));
    HIP_CHECK(hipStreamSynchronize(s));
    const double qmax = *std::max_element(Qp.begin(), Qp.end());
    std::vector<PrimPair> pp2;
    std::vector<PrimPairF> ppf2;
    std::vector<int> start2(ctx->nsp), count2(ctx->nsp);
    std::vector<int> keep;
    for (int sp = 0; sp < ctx->nsp; ++sp) {
      start2[sp] = (int)pp2.size();
      keep.clear();
      for (int t = ctx->pp_start[sp]; t < ctx->pp_start[sp] + ctx->pp_count[sp]; ++t)
        if (Qp[t] * qmax >= kPrimPairEps) keep.push_back(t);
      // largest bound first: the register kernels stop a primitive loop at the
      // first pair whose bound is too small (primitive-quartet screening)
      std::stable_sort(keep.begin(), keep.end(), [&](int x, int y) { return Qp[x] > Qp[y]; });
      for (int t : keep) {
        pp2.push_back(ctx->pp[t]);
        ppf2.push_back(ctx->ppf[t]);
        ppf2.back().q = Qp[t];
      }
      count2[sp] = (int)pp2.size() - start2[sp];
    }
    ctx->nprim_dropped = npp - (long long)pp2.size();
    ctx->pp_count0 = ctx->pp_count;
    {  // the screened, bound-sorted tables replace the first ones
      ctx->pp.swap(pp2);
      ctx->ppf.swap(ppf2);
      ctx->pp_start.swap(start2);
      ctx->pp_count.swap(count2);
      if ((rc = dev_upload(&d_pps, ctx->pp_start, own, s))) return rc;
      if ((rc = dev_upload(&d_ppc, ctx->pp_count, own, s))) return rc;
      if ((rc = dev_upload(&d_pp, ctx->pp, own, s))) return rc;
      if ((rc = dev_upload(&d_ppf, ctx->ppf, own, s))) return rc;
      dv.pp_start = d_pps;
      dv.pp_count = d_ppc;
      dv.pp = d_pp;
      dv.ppf = d_ppf;
    }
  }

  {  // the FP32 path's copy of the final primitive-pair table
    std::vector<PrimPairF32> f32(ctx->ppf.size());
    for (size_t t = 0; t < ctx->ppf.size(); ++t) {
      const PrimPairF& f = ctx->ppf[t];
      PrimPairF32& g = f32[t];
      g.Px = f.Px;
      g.Py = f.Py;
      g.Pz = f.Pz;
      g.eta = (float)f.eta;
      g.inv_eta = (float)f.inv_eta;
      g.ck = (float)f.ck;
      g.PAx = (float)f.PAx;
      g.PAy = (float)f.PAy;
      g.PAz = (float)f.PAz;
      g.q = (float)std::min(f.q, 3.0e38);
      g.pad = 0.0f;
    }
    if ((rc = dev_upload(&ctx->d_ppf32, f32, own, s))) return rc;
    dv.ppf32 = ctx->d_ppf32;
  }
  if (ctx->pp_count0.empty()) ctx->pp_count0 = ctx->pp_count;
  {
    int* d_ppc0;
    if ((rc = dev_upload(&d_ppc0, ctx->pp_count0, own, s))) return rc;
    dv.pp_count0 = d_ppc0;
  }

  // ---- Schwarz bounds (device) ----
  ctx->Q.assign(ctx->nsp, 1.0);
  if (ctx->eps_schwarz > 0.0) {
    std::vector<int> diag(ctx->nsp);
    std::iota(diag.begin(), diag.end(), 0);
    int* d_diag;
    if ((rc = dev_upload(&d_diag, diag, own, s))) return rc;
    ClassWork w{};
    w.la = w.lb = w.lc = w.ld = ctx->max_l;
    w.nwork = ctx->nsp;
    w.q_ab = d_diag;
    w.q_cd = d_diag;
    EriOut o{};
    o.mode = MODE_SCHWARZ;
    o.Qout = ctx->d_Q;
    HIP_CHECK(launch_eri_generic(dv, w, o, s));
    HIP_CHECK(hipMemcpyAsync(ctx->Q.data(), ctx->d_Q, sizeof(double) * ctx->nsp, hipMemcpyDeviceToHost, s));
    HIP_CHECK(hipStreamSynchronize(s));
  } else {
    HIP_CHECK(hipMemcpy(ctx->d_Q, ctx->Q.data(), sizeof(double) * ctx->nsp, hipMemcpyHostToDevice));
  }

  // ---- class lists: pairs of each class, Q descending (stable), Q > 0 ----
  // (per molecule of a batch: cls_pairs[m * ncls + class])
  const int ncls = pair_class(ASUKA_LMAX, ASUKA_LMAX) + 1;
  std::vector<std::vector<int>> cls_pairs((size_t)nmol * ncls);
  for (int sp = 0; sp < ctx->nsp; ++sp) {
    if (!(ctx->Q[sp] > 0.0)) continue;
    cls_pairs[(size_t)ctx->shell_mol[ctx->sp_A[sp]] * ncls +
              pair_class(ctx->shell_l[ctx->sp_A[sp]], ctx->shell_l[ctx->sp_B[sp]])]
        .push_back(sp);
  }
  for (auto& v : cls_pairs)
    std::stable_sort(v.begin(), v.end(), [&](int a, int b) { return ctx->Q[a] > ctx->Q[b]; });
  // rank of each shell pair in the stable Q-descending order of all pairs
  // (asuka/hf/direct_jk.py:1228 argsort(-Q, kind="stable")); inside a class
  // list the ranks increase
  std::vector<int> rank(ctx->nsp);
  {
    std::vector<int> ord(ctx->nsp);
    std::iota(ord.begin(), ord.end(), 0);
    std::stable_sort(ord.begin(), ord.end(), [&](int a, int b) { return ctx->Q[a] > ctx->Q[b]; });
    for (int r = 0; r < ctx->nsp; ++r) rank[ord[r]] = r;
  }
  std::vector<int> cls_la(ncls), cls_lb(ncls);
  for (int la = 0; la <= ASUKA_LMAX; ++la)
    for (int lb = 0; lb <= la; ++lb) {
      cls_la[pair_class(la, lb)] = la;
      cls_lb[pair_class(la, lb)] = lb;
    }
  // Pairs of each class are grouped by contraction depth (primitive pairs per
  // shell pair, in buckets no wider than 25 %), keeping the Q-descending order
  // inside a group and each pair's rank in the full class list. The kets of a
  // wavefront then have primitive loops of nearly equal length (each lane runs
  // its own); per bra pair its kets are a prefix of the ket group:
  // Q_ket >= eps / Q_bra and, within one class, rank_ket <= rank_bra (each
  // unordered pair of shell pairs once, asuka/hf/direct_jk.py:1228-1246).
  auto kbucket = [](int k) {
    int b = 1;
    while (b < k) b = std::max(b + 1, (b * 5) / 4);
    return b;
  };
  struct PairGroup {
    int kcount;  // bucket (largest primitive-pair count of the group's pairs)
    std::vector<int> pairs;  // Q-descending
    std::vector<int> ranks;  // position in the full class list
    std::vector<long long> kpref;  // prefix sums of the pairs' primitive-pair counts
    std::vector<long long> kpref0;  // the same before primitive-pair screening
    int col_start = 0;
  };
  std::vector<std::vector<PairGroup>> groups((size_t)nmol * ncls);
  std::vector<int> col_all, row_pair_all;
  std::vector<long long> row_off_all;
  std::vector<Segment> segs;
  std::vector<WaveUnit> units;
  std::vector<RowTask> tasks;
  for (size_t c = 0; c < cls_pairs.size(); ++c) {
    std::vector<int> kvals;
    for (int sp : cls_pairs[c]) kvals.push_back(kbucket(ctx->pp_count[sp]));
    std::sort(kvals.begin(), kvals.end());
    kvals.erase(std::unique(kvals.begin(), kvals.end()), kvals.end());
    for (int kv : kvals) {
      PairGroup g;
      g.kcount = kv;
      for (size_t r = 0; r < cls_pairs[c].size(); ++r)
        if (kbucket(ctx->pp_count[cls_pairs[c][r]]) == kv) {
          g.pairs.push_back(cls_pairs[c][r]);
          g.ranks.push_back((int)r);
        }
      g.kpref.assign(1, 0);
      for (int sp : g.pairs) g.kpref.push_back(g.kpref.back() + ctx->pp_count[sp]);
      g.kpref0.assign(1, 0);
      for (int sp : g.pairs) g.kpref0.push_back(g.kpref0.back() + ctx->pp_count0[sp]);
      // every group starts a 64-position block (wave units then start at
      // block boundaries, EriOut::jblk); -1 marks the padding
      col_all.resize((col_all.size() + 63) / 64 * 64, -1);
      g.col_start = (int)col_all.size();
      col_all.insert(col_all.end(), g.pairs.begin(), g.pairs.end());
      groups[c].push_back(std::move(g));
    }
  }
  const double eps = ctx->eps_schwarz;
  for (int X = 0; X < ncls; ++X) {
    for (int Y = 0; Y <= X; ++Y) {
      ClassPairHost cp{};
      cp.X = X;
      cp.Y = Y;
      cp.la = cls_la[X];
      cp.lb = cls_lb[X];
      cp.lc = cls_la[Y];
      cp.ld = cls_lb[Y];
      cp.seg_start = (int)segs.size();
      cp.unit_start = (long long)units.size();
      cp.task_start = (long long)tasks.size();
      const int nroots = (cp.la + cp.lb + cp.lc + cp.ld) / 2 + 1;
      int task_units = kTaskUnits;
      {  // LDS row accumulators and workgroup size of this class pair
        const int cC = ctx->nlao[cp.lc], cD = (cp.lc == cp.ld) ? 0 : ctx->nlao[cp.ld];
        const size_t bytes = sizeof(double) * (size_t)(ncart_c(cp.la) * ncart_c(cp.lb) +
                                                       (ncart_c(cp.la) + ncart_c(cp.lb)) * (cC + cD));
        const bool fits = bytes <= std::min(ctx->row_acc_max, ctx->lds_per_group - root_table_lds(nroots) - 1024);
        cp.acc_colsC = fits ? cC : 0;
        cp.acc_colsD = fits ? cD : 0;
        cp.wpg = (fits && bytes > kRowAccSmallBytes) ? 8 : 4;
        cp.task_units = task_units;
      }
      long long padded = 0;
      for (int m = 0; m < nmol; ++m)
      for (const PairGroup& bg : groups[(size_t)m * ncls + X]) {
        for (const PairGroup& kg : groups[(size_t)m * ncls + Y]) {
          const auto& cols = kg.pairs;
          long long off = 0;
          std::vector<long long> offs;
          std::vector<int> rowp;
          for (size_t r = 0; r < bg.pairs.size(); ++r) {
            const int P = bg.pairs[r];
            long long len;
            if (eps > 0.0) {
              // the reference's rule, Q_j >= eps / Q_i for the pair i of the
              // later rank in the global Q-descending order; the kets ranked
              // before P come first in cols
              const double thr = eps / std::max(ctx->Q[P], 1e-300);
              auto keep_before = [&](int c) { return ctx->Q[c] >= thr; };
              if (X == Y) {
                len = std::partition_point(cols.begin(), cols.end(), keep_before) - cols.begin();
              } else {
                const auto split = std::partition_point(cols.begin(), cols.end(),
                                                        [&](int c) { return rank[c] < rank[P]; });
                len = std::partition_point(cols.begin(), split, keep_before) - cols.begin();
                if (cols.begin() + len == split)
                  len = std::partition_point(split, cols.end(), [&](int c) {
                          return ctx->Q[P] >= eps / std::max(ctx->Q[c], 1e-300);
                        }) - cols.begin();
              }
            } else {
              len = (long long)cols.size();
            }
            if (X == Y)
              len = std::min<long long>(len, std::upper_bound(kg.ranks.begin(), kg.ranks.end(), bg.ranks[r]) -
                                                 kg.ranks.begin());
            if (len <= 0) continue;
            rowp.push_back(P);
            offs.push_back(off);
            off += len;
            {
              // the row's wave units (cut into row tasks below)
              const int nu = (int)((len + 63) / 64);
              cp.rows.emplace_back((int)(units.size() - cp.unit_start), nu);
              for (long long c0 = 0; c0 < len; c0 += 64) {
                const int nc = (int)std::min<long long>(64, len - c0);
                int nabu = ctx->pp_count[P];
                if (ctx->eps_schwarz > 0.0) {  // primitive-quartet cut of the bra loop (pairs sorted by bound)
                  double qw = 0.0;
                  for (int t = 0; t < nc; ++t) {
                    const int kq = kg.pairs[c0 + t];
                    if (ctx->pp_count[kq] > 0) qw = std::max(qw, ctx->ppf[ctx->pp_start[kq]].q);
                  }
                  const double thr = qw > 0.0 ? ctx->eps_prim / qw : 1e308;
                  int k = 0;
                  while (k < nabu && ctx->ppf[ctx->pp_start[P] + k].q >= thr) ++k;
                  nabu = k;
                }
                units.push_back(WaveUnit{P, kg.col_start + (int)c0, nc, nabu});
              }
            }
          }
          if (off == 0) continue;
          Segment sg{};
          sg.off = padded;
          sg.len = off;
          sg.row_start = (int)row_off_all.size();
          sg.nrows = (int)offs.size();
          sg.col_start = kg.col_start;
          sg.kab = bg.kcount;
          sg.kcd = kg.kcount;
          segs.push_back(sg);
          row_pair_all.insert(row_pair_all.end(), rowp.begin(), rowp.end());
          row_pair_all.push_back(-1);  // keeps row_pair aligned with row_off's trailing total
          row_off_all.insert(row_off_all.end(), offs.begin(), offs.end());
          row_off_all.push_back(off);
          padded += (off + 63) / 64 * 64;
          cp.nquartets += off;
          for (size_t r = 0; r < rowp.size(); ++r) {  // primitive quartets evaluated
            const long long len = (r + 1 < rowp.size() ? offs[r + 1] : off) - offs[r];
            cp.nprim_quartets += (long long)ctx->pp_count[rowp[r]] * (kg.kpref[len] - kg.kpref[0]);
            cp.nprim_quartets0 += (long long)ctx->pp_count0[rowp[r]] * (kg.kpref0[len] - kg.kpref0[0]);
          }
        }
      }
      cp.nseg = (int)segs.size() - cp.seg_start;
      cp.nunits = (long long)units.size() - cp.unit_start;
      {
        const std::vector<RowTask> tk = make_tasks(cp, units.data() + cp.unit_start, col_all, ctx->pp_count, task_units);
        tasks.insert(tasks.end(), tk.begin(), tk.end());
        cp.ntasks = cp.ntasks0 = (long long)tk.size();
        cp.task_units0 = task_units;
      }

      if (cp.nseg == 0) continue;
      cp.nwork = padded;
      cp.nprim_roots = cp.nprim_quartets * nroots;
      {  // EFM-v1 per primitive quartet and per quartet (digestion)
        const int nmax = cp.la + cp.lb, mmax = cp.lc + cp.ld;
        const double n1 = (double)(cp.la + 1) * (cp.lb + 1) * (cp.lc + 1) * (cp.ld + 1);
        const double ncomp = (double)ncart_c(cp.la) * ncart_c(cp.lb) * ncart_c(cp.lc) * ncart_c(cp.ld);
        const double per_prim = 24.0 + nroots * (24.0 + 15.0 * (nmax + 1) * (mmax + 1) + 6.0 * n1 + 4.0 * ncomp);
        cp.cost = per_prim * (double)cp.nprim_quartets + 8.0 * ncomp * (double)cp.nquartets;
      }
      ctx->cps.push_back(cp);
      ctx->ntasks += cp.nquartets;
    }
  }
  assign_streams(ctx);
  col_all.resize((col_all.size() + 63) / 64 * 64, -1);
  if ((rc = pos_layout(ctx, col_all))) return rc;
  if ((rc = dev_upload(&ctx->d_segs, segs, own, s))) return rc;
  if ((rc = dev_upload(&ctx->d_units, units, own, s))) return rc;
  if ((rc = dev_upload(&ctx->d_tasks, tasks, own, s))) return rc;
  if ((rc = dev_upload(&ctx->d_row_pair, row_pair_all, own, s))) return rc;
  if ((rc = dev_upload(&ctx->d_row_off, row_off_all, own, s))) return rc;
  if ((rc = dev_upload(&ctx->d_col_pair, col_all, own, s))) return rc;
  ctx->h_units = std::move(units);  // kept for re-cutting row tasks (autotuner)
  ctx->h_col = std::move(col_all);
  {
    std::vector<unsigned long long> cz(kCountStride * std::max<size_t>(1, ctx->cps.size()), 0ull);
    if ((rc = dev_upload(&ctx->d_counters, cz, own, s))) return rc;
    // two work-queue heads per class pair: the FP64 register kernel's, then the FP32 one's
    std::vector<unsigned int> qz(2 * std::max<size_t>(1, ctx->cps.size()), 0u);
    if ((rc = dev_upload(&ctx->d_queue, qz, own, s))) return rc;
  }
  if ((rc = route_alloc(ctx))) return rc;
  HIP_CHECK(hipStreamSynchronize(s));
  *out = guard.release();
  return ASUKA_OK;
}

extern "C" {

int asuka_jk_ctx_create(const asuka_basis_desc* bd, const asuka_jk_opts* opts, void* stream, asuka_jk_ctx** out) {
  return ctx_build(bd, nullptr, opts, (hipStream_t)stream, out);
}

int asuka_jk_batch_create(int nmol, const asuka_basis_desc* const* bases, const asuka_jk_opts* opts, void* stream,
                          asuka_jk_ctx** out) {
  if (nmol < 1 || !bases || !opts || !out) return fail(ASUKA_E_INVALID, "nmol must be >= 1 and bases, opts, out non-null");
  *out = nullptr;
  // one basis holding every molecule's shells, AOs and primitives back to back
  std::vector<double> cxyz, pexp, pcoef, axyz, aZ;
  std::vector<int> pst, pnp, sl, sao, sat, smol;
  int nao = 0, nprim = 0, natm = 0;
  bool atoms = true;
  for (int m = 0; m < nmol; ++m) {
    const asuka_basis_desc* b = bases[m];
    if (!b || b->nshell <= 0 || !b->shell_cxyz || !b->shell_prim_start || !b->shell_nprim || !b->shell_l ||
        !b->shell_ao_start || !b->prim_exp || !b->prim_coef)
      return fail(ASUKA_E_INVALID, "every basis must be non-null with non-null arrays and nshell > 0");
    int mnao = 0;
    for (int i = 0; i < b->nshell; ++i) {
      if (b->shell_l[i] < 0 || b->shell_l[i] > ASUKA_LMAX)
        return fail(b->shell_l[i] < 0 ? ASUKA_E_INVALID : ASUKA_E_UNSUPPORTED,
                    "direct J/K supports only 0 <= l <= " + std::to_string(ASUKA_LMAX));
      cxyz.insert(cxyz.end(), b->shell_cxyz + 3 * i, b->shell_cxyz + 3 * i + 3);
      pst.push_back(nprim + b->shell_prim_start[i]);
      pnp.push_back(b->shell_nprim[i]);
      sl.push_back(b->shell_l[i]);
      sao.push_back(nao + b->shell_ao_start[i]);
      sat.push_back(natm + (b->shell_atom ? b->shell_atom[i] : 0));
      smol.push_back(m);
      mnao = std::max(mnao, b->shell_ao_start[i] + ncart_c(b->shell_l[i]));
    }
    pexp.insert(pexp.end(), b->prim_exp, b->prim_exp + b->nprim);
    pcoef.insert(pcoef.end(), b->prim_coef, b->prim_coef + b->nprim);
    if (b->natm > 0 && b->atom_xyz && b->atom_charge) {
      axyz.insert(axyz.end(), b->atom_xyz, b->atom_xyz + 3 * b->natm);
      aZ.insert(aZ.end(), b->atom_charge, b->atom_charge + b->natm);
    } else {
      atoms = false;
    }
    nao += mnao;
    nprim += b->nprim;
    natm += std::max(0, b->natm);
  }
  asuka_basis_desc c{};
  c.nshell = (int32_t)sl.size();
  c.shell_cxyz = cxyz.data();
  c.shell_prim_start = pst.data();
  c.shell_nprim = pnp.data();
  c.shell_l = sl.data();
  c.shell_ao_start = sao.data();
  c.shell_atom = sat.data();
  c.nprim = nprim;
  c.prim_exp = pexp.data();
  c.prim_coef = pcoef.data();
  c.natm = atoms ? natm : 0;
  c.atom_xyz = atoms ? axyz.data() : nullptr;
  c.atom_charge = atoms ? aZ.data() : nullptr;
  asuka_jk_ctx* ctx = nullptr;
  int rc = ctx_build(&c, &smol, opts, (hipStream_t)stream, &ctx);
  if (rc) return rc;
  std::unique_ptr<asuka_jk_ctx> guard(ctx);
  for (double** p : {&ctx->d_Dbig, &ctx->d_Jbig, &ctx->d_Kbig}) {
    const hipError_t e = hipMalloc((void**)p, sizeof(double) * (size_t)ctx->nelem);
    if (e != hipSuccess) return fail(ASUKA_E_NOMEM, std::string("hipMalloc: ") + hipGetErrorString(e));
    ctx->owned.push_back(*p);
  }
  *out = guard.release();
  return ASUKA_OK;
}

int asuka_jk_batch_info(const asuka_jk_ctx* ctx, int32_t* mol_nao, int cap) {
  if (!ctx || (cap > 0 && !mol_nao)) return fail(ASUKA_E_INVALID, "ctx and mol_nao must be non-null");
  for (int m = 0; m < std::min(cap, ctx->nmol); ++m) mol_nao[m] = ctx->mol_nao[m];
  return ctx->nmol;
}

int asuka_jk_ctx_destroy(asuka_jk_ctx* ctx) {
  if (!ctx) return ASUKA_OK;
  (void)hipDeviceSynchronize();
  delete ctx;
  return ASUKA_OK;
}

int asuka_jk_ctx_info(const asuka_jk_ctx* ctx, int64_t out[8]) {
  if (!ctx || !out) return fail(ASUKA_E_INVALID, "ctx and out must be non-null");
  out[0] = ctx->nao;
  out[1] = ctx->nshell;
  out[2] = ctx->nsp;
  out[3] = ctx->ntasks;
  out[4] = ctx->max_l;
  out[5] = (int64_t)ctx->cps.size();
  out[6] = ctx->natm;
  out[7] = (int64_t)ctx->pp.size();
  return ASUKA_OK;
}

int asuka_jk_shell_pairs(const asuka_jk_ctx* ctx, int32_t* sp_A, int32_t* sp_B, double* Q) {
  if (!ctx) return fail(ASUKA_E_INVALID, "ctx must be non-null");
  for (int i = 0; i < ctx->nsp; ++i) {
    if (sp_A) sp_A[i] = ctx->sp_A[i];
    if (sp_B) sp_B[i] = ctx->sp_B[i];
    if (Q) Q[i] = ctx->Q[i];
  }
  return ASUKA_OK;
}

int asuka_jk_class_stats(const asuka_jk_ctx* ctx, int64_t* rows, int cap) {
  if (!ctx) return fail(ASUKA_E_INVALID, "ctx must be non-null");
  int n = 0;
  for (const auto& cp : ctx->cps) {
    if (n >= cap) break;
    int64_t* r = rows + 8 * n;
    r[0] = cp.la;
    r[1] = cp.lb;
    r[2] = cp.lc;
    r[3] = cp.ld;
    r[4] = cp.nquartets;
    r[5] = cp.nprim_quartets;
    r[6] = cp.nprim_roots;
    r[7] = cp.nprim_quartets0;
    ++n;
  }
  return n;
}

}  // extern "C"

static ClassWork work_of(const asuka_jk_ctx* ctx, const ClassPairHost& cp, int index) {
  ClassWork w{};
  w.cp_index = index;
  w.la = cp.la;
  w.lb = cp.lb;
  w.lc = cp.lc;
  w.ld = cp.ld;
  w.seg = ctx->d_segs + cp.seg_start;
  w.nseg = cp.nseg;
  w.row_pair = ctx->d_row_pair;
  w.row_off = ctx->d_row_off;
  w.col_pair = ctx->d_col_pair;
  w.nwork = cp.nwork;
  w.units = ctx->d_units + cp.unit_start;
  w.nunits = cp.nunits;
  w.tasks = cp.d_tasks_own ? cp.d_tasks_own : ctx->d_tasks + cp.task_start;
  w.ntasks = cp.ntasks;
  w.acc_colsC = cp.acc_colsC;
  w.acc_colsD = cp.acc_colsD;
  w.lao_offC = ctx->lao_off[cp.lc];
  w.lao_offD = ctx->lao_off[cp.ld];
  w.queue = ctx->d_queue + index;
  w.queue32 = ctx->d_queue + std::max<size_t>(1, ctx->cps.size()) + index;
  w.wpg = cp.wpg;
  return w;
}

static inline unsigned grid_for(long long n, int block) {
  long long g = (n + block - 1) / block;
  return (unsigned)std::max<long long>(1, std::min<long long>(g, 65535));
}

// Fork: the context's side streams wait for everything issued on s so far.
static int fork_streams(asuka_jk_ctx* ctx, hipStream_t s) {
  if (ctx->side.empty()) return ASUKA_OK;
  HIP_CHECK(hipEventRecord(ctx->ev_fork, s));
  for (hipStream_t st : ctx->side) HIP_CHECK(hipStreamWaitEvent(st, ctx->ev_fork, 0));
  return ASUKA_OK;
}

// Join: s waits for everything issued on the side streams.
static int join_streams(asuka_jk_ctx* ctx, hipStream_t s) {
  for (size_t k = 0; k < ctx->side.size(); ++k) {
    HIP_CHECK(hipEventRecord(ctx->ev_join[k], ctx->side[k]));
    HIP_CHECK(hipStreamWaitEvent(s, ctx->ev_join[k], 0));
  }
  return ASUKA_OK;
}

// The stream class pair cp is launched on.
static hipStream_t stream_of(const asuka_jk_ctx* ctx, const ClassPairHost& cp, hipStream_t s) {
  return (cp.stream > 0 && cp.stream <= (int)ctx->side.size()) ? ctx->side[cp.stream - 1] : s;
}

// queue: the class pairs' work-queue heads (the context's, or the static
// scratch of asuka_jk_build_noalloc); dev: the device tables the kernels read
static int run_classes(asuka_jk_ctx* ctx, const EriOut& o, hipStream_t s, bool timed, unsigned int* queue = nullptr,
                       const DevBasis* dev = nullptr) {
  if (!queue) queue = ctx->d_queue;
  if (!dev) dev = &ctx->dev;
  if (timed) {
    if (!ctx->ev_start) HIP_CHECK(hipEventCreate(&ctx->ev_start));
    if (!ctx->ev_stop) HIP_CHECK(hipEventCreate(&ctx->ev_stop));
    HIP_CHECK(hipEventRecord(ctx->ev_start, s));
  }
  const long long ncp = (long long)std::max<size_t>(1, ctx->cps.size());
  {  // work-queue heads start at zero (a kernel, not a memset: the build may be graph-captured)
    const long long nq = 2 * ncp;
    hipLaunchKernelGGL(k_zero_u32, dim3((unsigned)((nq + 255) / 256)), dim3(256), 0, s, queue, nq);
    HIP_CHECK(hipGetLastError());
  }
  // Mixed builds: the precision vote of every wave unit once, before the
  // launches, and the units split (in order) into an FP64 and an FP32 list, so
  // each kernel of a class pair walks only its own units.
  const long long nu_all = (long long)ctx->h_units.size();
  const bool routed = o.mixed && ctx->d_rf64 && nu_all > 0;
  if (routed) {
    hipLaunchKernelGGL(k_pos_dmax, dim3((unsigned)ctx->nblk), dim3(64), 0, s, ctx->d_col_pair, dev->Dsp,
                       ctx->d_dmaxcol);
    HIP_CHECK(hipGetLastError());
    hipLaunchKernelGGL(k_route_units, dim3(grid_for(nu_all, 256)), dim3(256), 0, s, *dev, ctx->d_units, ctx->d_qcol,
                       ctx->d_dmaxcol, nu_all, o.eps_density, o.tau_mixed, ctx->d_rf64, ctx->d_rf32);
    HIP_CHECK(hipGetLastError());
    size_t tb = ctx->scan_tmp_bytes;
    HIP_CHECK(hipcub::DeviceScan::ExclusiveSum(ctx->d_scan_tmp, tb, ctx->d_rf64, ctx->d_rp64, (int)(nu_all + 1), s));
    tb = ctx->scan_tmp_bytes;
    HIP_CHECK(hipcub::DeviceScan::ExclusiveSum(ctx->d_scan_tmp, tb, ctx->d_rf32, ctx->d_rp32, (int)(nu_all + 1), s));
    hipLaunchKernelGGL(k_scatter_units, dim3(grid_for(nu_all, 256)), dim3(256), 0, s, ctx->d_units, nu_all,
                       ctx->d_rf64, ctx->d_rf32, ctx->d_rp64, ctx->d_rp32, ctx->d_units64, ctx->d_units32);
    HIP_CHECK(hipGetLastError());
  }
  // J[c,d] and D[c,d] by ket position for the register kernels of J/K and
  // Fock builds (EriOut::jblk)
  EriOut oo = o;
  double* Jt = (o.mode == MODE_FOCK) ? o.F : (o.mode == MODE_JK ? o.J : nullptr);
  double* Jt2 = (o.mode == MODE_JK && o.D2) ? o.J2 : nullptr;
  const bool pos = ctx->d_jblk && ctx->nblk > 0 && (Jt || Jt2);
  if (pos) {
    oo.jblk = ctx->d_jblk;
    oo.Jpos = ctx->d_jpos;
    oo.Dpos = ctx->d_dpos;
    if (o.D2) {
      oo.Jpos2 = ctx->d_jpos + ctx->pos_len;
      oo.Dpos2 = ctx->d_dpos + ctx->pos_len;
    }
    hipLaunchKernelGGL(k_pos_gather, dim3((unsigned)ctx->nblk), dim3(64), 0, s, *dev, ctx->d_col_pair, ctx->d_jblk,
                       o.D, o.D2, ctx->d_dpos, ctx->d_jpos, ctx->pos_len);
    HIP_CHECK(hipGetLastError());
  }
  // K rows of the row-task flush with permuted columns (EriOut::Kc)
  double* Kt = (o.mode == MODE_FOCK) ? o.F : (o.mode == MODE_JK ? o.K : nullptr);
  double* Kt2 = (o.mode == MODE_JK && o.D2) ? o.K2 : nullptr;
  const long long nn = (long long)ctx->nao * ctx->nao;
  const bool perm = ctx->d_kc && (Kt || Kt2);
  if (perm) {
    hipLaunchKernelGGL(k_fill, dim3(grid_for((Kt2 ? 2 : 1) * nn, 256)), dim3(256), 0, s, ctx->d_kc,
                       (Kt2 ? 2 : 1) * nn, 0.0);
    HIP_CHECK(hipGetLastError());
    if (Kt) oo.Kc = ctx->d_kc;
    if (Kt2) oo.Kc2 = ctx->d_kc + nn;
  }
  int rc = fork_streams(ctx, s);
  if (rc) return rc;
  long long n = 0;
  for (int i : ctx->launch_order) {
    const ClassPairHost& cph = ctx->cps[i];
    hipStream_t ls = stream_of(ctx, cph, s);
    ClassWork w = work_of(ctx, cph, i);
    w.queue = queue + i;
    w.queue32 = queue + ncp + i;
    if (routed && w.ntasks > 0 && eri_fast_has_plan(w.la, w.lb, w.lc, w.ld)) {
      RowTask* t64 = ctx->d_rtasks64 + cph.rtask_off;
      RowTask* t32 = ctx->d_rtasks32 + cph.rtask_off;
      hipLaunchKernelGGL(k_route_tasks, dim3(grid_for(w.ntasks, 256)), dim3(256), 0, ls, w.tasks, w.ntasks,
                         cph.unit_start, ctx->d_rp64, ctx->d_rp32, t64, t32);
      HIP_CHECK(hipGetLastError());
      w.routed = 1;
      w.units = ctx->d_units64;
      w.tasks = t64;
      w.units32 = ctx->d_units32;
      w.tasks32 = t32;
    }
    if (o.D2 && w.acc_colsC > 0) {  // two accumulator sets must fit the workgroup's LDS
      const size_t nA = ncart_c(w.la), nB = ncart_c(w.lb);
      const size_t bytes = 2 * sizeof(double) * (nA * nB + (nA + nB) * (size_t)(w.acc_colsC + w.acc_colsD));
      if (bytes > ctx->lds_per_group - root_table_lds((w.la + w.lb + w.lc + w.ld) / 2 + 1) - 1024)
        w.acc_colsC = w.acc_colsD = 0;
    }
    bool handled = false;
    HIP_CHECK(launch_eri_fast(*dev, w, oo, ls, &handled));
    if (!handled) HIP_CHECK(launch_eri_coop(*dev, w, o, ls, &handled));
    if (!handled) HIP_CHECK(launch_eri_generic(*dev, w, o, ls));
    ++n;
  }
  if ((rc = join_streams(ctx, s))) return rc;
  if (pos) {
    hipLaunchKernelGGL(k_pos_scatter, dim3((unsigned)ctx->nblk), dim3(64), 0, s, *dev, ctx->d_col_pair, ctx->d_jblk,
                       ctx->d_jpos, Jt, Jt2, ctx->pos_len, o.fixed);
    HIP_CHECK(hipGetLastError());
  }
  if (perm) {
    if (Kt)
      hipLaunchKernelGGL(k_unperm_add, dim3(grid_for(nn, 256)), dim3(256), 0, s, Kt, ctx->d_kc, ctx->d_ao_g, ctx->nao,
                         o.fixed);
    if (Kt2)
      hipLaunchKernelGGL(k_unperm_add, dim3(grid_for(nn, 256)), dim3(256), 0, s, Kt2, ctx->d_kc + nn, ctx->d_ao_g,
                         ctx->nao, o.fixed);
    HIP_CHECK(hipGetLastError());
  }
  if (timed) HIP_CHECK(hipEventRecord(ctx->ev_stop, s));
  ctx->last_launches = n;
  return ASUKA_OK;
}



// J/K of one density (D2 == nullptr) or of two densities from one evaluation
// of each quartet; the outputs are zeroed first and symmetrised at the end.
// queue, dsp (both or neither): per-build scratch in place of the context's
// work-queue heads and shell-pair density maxima (asuka_jk_build_noalloc).
static int jk_build_impl(asuka_jk_ctx* ctx, const double* D, const double* D2, double* J, double* K, double* J2,
                         double* K2, bool wantJ, bool wantK, int flags, hipStream_t s, bool symmetrize = true,
                         unsigned int* queue = nullptr, double* dsp = nullptr) {
  const long long nn = ctx->nelem;
  double* outs[4] = {wantJ ? J : nullptr, wantK ? K : nullptr, (D2 && wantJ) ? J2 : nullptr,
                     (D2 && wantK) ? K2 : nullptr};
  for (double* X : outs)
    if (X) hipLaunchKernelGGL(k_fill, dim3(grid_for(nn, 256)), dim3(256), 0, s, X, nn, 0.0);
  DevBasis dev = ctx->dev;
  if (dsp) dev.Dsp = dsp;
  hipLaunchKernelGGL(k_dsp, dim3(grid_for(ctx->nsp, 128)), dim3(128), 0, s, D, D2, dev, (double*)dev.Dsp, ctx->d_kd,
                     ctx->d_pos_of_sp);
  HIP_CHECK(hipGetLastError());
  EriOut o{};
  o.mode = MODE_JK;
  o.D = D;
  o.J = outs[0];
  o.K = outs[1];
  if (D2) {
    o.D2 = D2;
    o.J2 = outs[2];
    o.K2 = outs[3];
  }
  o.eps_density = ctx->eps_density;
  o.eps_prim = ctx->eps_schwarz > 0.0 ? ctx->eps_prim : 0.0;
  o.fixed = (flags & ASUKA_JK_DETERMINISTIC) ? 1 : 0;  // the zeroed outputs are fixed-point zeros too
  o.mixed = ((flags & ASUKA_JK_MIXED) && !o.fixed) ? 1 : 0;  // deterministic builds run every quartet in FP64
  o.tau_mixed = ctx->tau_mixed;
  if (flags & ASUKA_JK_COUNT) {
    const long long nc = kCountStride * (long long)std::max<size_t>(1, ctx->cps.size());
    hipLaunchKernelGGL(k_zero_u64, dim3(grid_for(nc, 256)), dim3(256), 0, s, ctx->d_counters, nc);
    o.counters = ctx->d_counters;
  }
  int rc = run_classes(ctx, o, s, (flags & ASUKA_JK_TIME) != 0, queue, &dev);
  if (rc) return rc;
  if (o.fixed) {
    for (double* X : outs)
      if (X) hipLaunchKernelGGL(k_from_fixed, dim3(grid_for(nn, 256)), dim3(256), 0, s, X, nn);
    HIP_CHECK(hipGetLastError());
  }
  if (!symmetrize) return ASUKA_OK;
  const int n = ctx->nao;
  dim3 blk(16, 16), grd((n + 15) / 16, (n + 15) / 16);
  for (double* X : outs)
    if (X) hipLaunchKernelGGL(k_symmetrize, grd, blk, 0, s, X, n);
  HIP_CHECK(hipGetLastError());
  return ASUKA_OK;
}

static int reject_batch(const asuka_jk_ctx* ctx) {
  return ctx->nmol > 1 ? fail(ASUKA_E_INVALID, "a batch context must be used with asuka_jk_batch_build") : ASUKA_OK;
}

extern "C" {

int asuka_jk_batch_build(asuka_jk_ctx* ctx, const double* const* D, double* const* J, double* const* K, int flags,
                         void* stream) {
  if (!ctx || !D) return fail(ASUKA_E_INVALID, "ctx and D must be non-null");
  if (flags & ASUKA_JK_DETERMINISTIC) return fail(ASUKA_E_INVALID, "batch builds do not take ASUKA_JK_DETERMINISTIC");
  const int nmol = ctx->nmol;
  for (int m = 0; m < nmol; ++m)
    if (!D[m]) return fail(ASUKA_E_INVALID, "every D[m] must be non-null");
  const bool wantJ = (flags & ASUKA_JK_WANT_J) && J;
  const bool wantK = (flags & ASUKA_JK_WANT_K) && K;
  if (!wantJ && !wantK) return ASUKA_OK;
  hipStream_t s = (hipStream_t)stream;
  if (!ctx->d_Dbig) return fail(ASUKA_E_INVALID, "ctx is not a batch context (asuka_jk_batch_create)");
  long long base = 0;
  std::vector<long long> bases(nmol);
  for (int m = 0; m < nmol; ++m) {  // the molecules' densities into their packed blocks
    const long long nn = (long long)ctx->mol_nao[m] * ctx->mol_nao[m];
    bases[m] = base;
    hipLaunchKernelGGL(k_copy, dim3(grid_for(nn, 256)), dim3(256), 0, s, ctx->d_Dbig + base, D[m], nn);
    base += nn;
  }
  HIP_CHECK(hipGetLastError());
  int rc = jk_build_impl(ctx, ctx->d_Dbig, nullptr, wantJ ? ctx->d_Jbig : nullptr, wantK ? ctx->d_Kbig : nullptr,
                         nullptr, nullptr, wantJ, wantK, flags, s, false);
  if (rc) return rc;
  for (int m = 0; m < nmol; ++m) {  // symmetrised blocks to the caller's matrices
    const int n = ctx->mol_nao[m];
    const long long nn = (long long)n * n;
    if (wantJ && J[m]) hipLaunchKernelGGL(k_sym_copy, dim3(grid_for(nn, 256)), dim3(256), 0, s, J[m], ctx->d_Jbig + bases[m], n);
    if (wantK && K[m]) hipLaunchKernelGGL(k_sym_copy, dim3(grid_for(nn, 256)), dim3(256), 0, s, K[m], ctx->d_Kbig + bases[m], n);
  }
  HIP_CHECK(hipGetLastError());
  return ASUKA_OK;
}

int asuka_jk_build(asuka_jk_ctx* ctx, const double* D, double* J, double* K, int flags, void* stream) {
  if (!ctx || !D) return fail(ASUKA_E_INVALID, "ctx and D must be non-null");
  if (reject_batch(ctx)) return ASUKA_E_INVALID;
  const bool wantJ = (flags & ASUKA_JK_WANT_J) && J;
  const bool wantK = (flags & ASUKA_JK_WANT_K) && K;
  if (!wantJ && !wantK) return ASUKA_OK;
  return jk_build_impl(ctx, D, nullptr, J, K, nullptr, nullptr, wantJ, wantK, flags, (hipStream_t)stream);
}

}  // extern "C"

// Per-build scratch of asuka_jk_build_noalloc: the class pairs' work-queue
// heads, then the shell-pair density maxima. One copy per GPU, kept until
// the process ends; host code addresses it only through the pointer
// hipGetSymbolAddress returns on that device.
static constexpr size_t kNoAllocScratchBytes = 4u << 20;
static constexpr size_t kNoAllocQueueBytes = 16u << 10;  // 4096 work-queue heads (2048 class pairs)
static __device__ __align__(256) unsigned char scratch[kNoAllocScratchBytes];

// Device address of `scratch` on the current device, looked up once per device.
static int noalloc_scratch(unsigned char** out) {
  static std::mutex mu;
  static std::vector<unsigned char*> ptrs;
  int dev = 0;
  HIP_CHECK(hipGetDevice(&dev));
  std::lock_guard<std::mutex> lk(mu);
  if ((int)ptrs.size() <= dev) ptrs.resize(dev + 1, nullptr);
  if (!ptrs[dev]) {
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess) return fail(ASUKA_E_HIP, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
    ptrs[dev] = (unsigned char*)p;
  }
  *out = ptrs[dev];
  return ASUKA_OK;
}

extern "C" {

int asuka_jk_build_noalloc(asuka_jk_ctx* ctx, const double* D, double* J, double* K, int flags, void* stream) {
  if (!ctx || !D) return fail(ASUKA_E_INVALID, "ctx and D must be non-null");
  if (reject_batch(ctx)) return ASUKA_E_INVALID;
  const size_t nqueue = 2 * std::max<size_t>(1, ctx->cps.size());  // two heads per class pair
  if (nqueue * sizeof(unsigned int) > kNoAllocQueueBytes)
    return fail(ASUKA_E_INVALID, "asuka_jk_build_noalloc: " + std::to_string(ctx->cps.size()) +
                                     " class pairs exceed the scratch limit of " +
                                     std::to_string(kNoAllocQueueBytes / sizeof(unsigned int) / 2));
  if ((size_t)ctx->nsp * sizeof(double) > kNoAllocScratchBytes - kNoAllocQueueBytes)
    return fail(ASUKA_E_INVALID, "asuka_jk_build_noalloc: " + std::to_string(ctx->nsp) +
                                     " shell pairs exceed the scratch limit of " +
                                     std::to_string((kNoAllocScratchBytes - kNoAllocQueueBytes) / sizeof(double)));
  if (flags & ASUKA_JK_TIME) return fail(ASUKA_E_INVALID, "asuka_jk_build_noalloc does not take ASUKA_JK_TIME");
  const bool wantJ = (flags & ASUKA_JK_WANT_J) && J;
  const bool wantK = (flags & ASUKA_JK_WANT_K) && K;
  if (!wantJ && !wantK) return ASUKA_OK;
  unsigned char* base = nullptr;
  int rc = noalloc_scratch(&base);
  if (rc) return rc;
  return jk_build_impl(ctx, D, nullptr, J, K, nullptr, nullptr, wantJ, wantK, flags, (hipStream_t)stream, true,
                       (unsigned int*)base, (double*)(base + kNoAllocQueueBytes));
}

int asuka_jk_build_multi(asuka_jk_ctx* ctx, const double* Da, const double* Db, double* Ja, double* Ka, double* Jb,
                         double* Kb, int flags, void* stream) {
  if (!ctx || !Da || !Db) return fail(ASUKA_E_INVALID, "ctx, Da and Db must be non-null");
  if (reject_batch(ctx)) return ASUKA_E_INVALID;
  const bool wantJ = (flags & ASUKA_JK_WANT_J) != 0;
  const bool wantK = (flags & ASUKA_JK_WANT_K) != 0;
  if (wantJ && (!Ja || !Jb)) return fail(ASUKA_E_INVALID, "Ja and Jb must be non-null with ASUKA_JK_WANT_J");
  if (wantK && (!Ka || !Kb)) return fail(ASUKA_E_INVALID, "Ka and Kb must be non-null with ASUKA_JK_WANT_K");
  if (!wantJ && !wantK) return ASUKA_OK;
  const double* outs[4] = {Ja, Ka, Jb, Kb};
  for (int i = 0; i < 4; ++i)
    for (int j = 0; j < i; ++j)
      if (outs[i] && outs[i] == outs[j]) return fail(ASUKA_E_INVALID, "Ja, Ka, Jb and Kb must be distinct buffers");
  return jk_build_impl(ctx, Da, Db, Ja, Ka, Jb, Kb, wantJ, wantK, flags, (hipStream_t)stream);
}

int asuka_fock_build(asuka_jk_ctx* ctx, const double* D, const double* hcore, double* F, int flags, void* stream) {
  if (!ctx || !D || !hcore || !F) return fail(ASUKA_E_INVALID, "ctx, D, hcore and F must be non-null");
  if (reject_batch(ctx)) return ASUKA_E_INVALID;
  if (F == D) return fail(ASUKA_E_INVALID, "F must not alias D");
  hipStream_t s = (hipStream_t)stream;
  const long long nn = (long long)ctx->nao * ctx->nao;
  if (F != hcore) hipLaunchKernelGGL(k_copy, dim3(grid_for(nn, 256)), dim3(256), 0, s, F, hcore, nn);
  const bool fixed = (flags & ASUKA_JK_DETERMINISTIC) != 0;
  if (fixed) hipLaunchKernelGGL(k_to_fixed, dim3(grid_for(nn, 256)), dim3(256), 0, s, F, nn);
  hipLaunchKernelGGL(k_dsp, dim3(grid_for(ctx->nsp, 128)), dim3(128), 0, s, D, nullptr, ctx->dev, ctx->d_Dsp,
                     ctx->d_kd, ctx->d_pos_of_sp);
  HIP_CHECK(hipGetLastError());
  EriOut o{};
  o.mode = MODE_FOCK;
  o.fixed = fixed ? 1 : 0;
  o.D = D;
  o.F = F;
  o.eps_density = ctx->eps_density;
  o.eps_prim = ctx->eps_schwarz > 0.0 ? ctx->eps_prim : 0.0;
  o.mixed = ((flags & ASUKA_JK_MIXED) && !fixed) ? 1 : 0;  // deterministic builds run every quartet in FP64
  o.tau_mixed = ctx->tau_mixed;
  if (flags & ASUKA_JK_COUNT) {
    const long long nc = kCountStride * (long long)std::max<size_t>(1, ctx->cps.size());
    hipLaunchKernelGGL(k_zero_u64, dim3(grid_for(nc, 256)), dim3(256), 0, s, ctx->d_counters, nc);
    o.counters = ctx->d_counters;
  }
  int rc = run_classes(ctx, o, s, (flags & ASUKA_JK_TIME) != 0);
  if (rc) return rc;
  if (fixed) hipLaunchKernelGGL(k_from_fixed, dim3(grid_for(nn, 256)), dim3(256), 0, s, F, nn);
  const int n = ctx->nao;
  dim3 blk(16, 16), grd((n + 15) / 16, (n + 15) / 16);
  hipLaunchKernelGGL(k_symmetrize, grd, blk, 0, s, F, n);
  HIP_CHECK(hipGetLastError());
  return ASUKA_OK;
}

int asuka_jk_build_batched(int nmol, asuka_jk_ctx* const* ctxs, const double* const* D, double* const* J,
                           double* const* K, int flags, void* stream) {
  if (nmol < 0) return fail(ASUKA_E_INVALID, "nmol must be >= 0");
  if (flags & ASUKA_JK_DETERMINISTIC) return fail(ASUKA_E_INVALID, "batch builds do not take ASUKA_JK_DETERMINISTIC");
  if (nmol > 0 && (!ctxs || !D)) return fail(ASUKA_E_INVALID, "ctxs and D must be non-null");
  for (int m = 0; m < nmol; ++m)
    if (!ctxs[m] || !D[m]) return fail(ASUKA_E_INVALID, "every ctxs[m] and D[m] must be non-null");
  for (int m = 0; m < nmol; ++m) {
    const int rc = asuka_jk_build(ctxs[m], D[m], J ? J[m] : nullptr, K ? K[m] : nullptr, flags & ~ASUKA_JK_TIME,
                                  stream);
    if (rc) return rc;
  }
  return ASUKA_OK;
}

}  // extern "C"

// Row tasks of class pair i re-cut to task_units units (device copy owned by
// the context; the original cut stays in d_tasks).
static int recut_tasks(asuka_jk_ctx* ctx, int i, int task_units, hipStream_t s) {
  ClassPairHost& cp = ctx->cps[i];
  if (task_units == cp.task_units) return ASUKA_OK;
  cp.task_units = task_units;
  if (task_units == cp.task_units0) {
    cp.d_tasks_own = nullptr;
    cp.ntasks = cp.ntasks0;
    return ASUKA_OK;
  }
  for (const auto& c : cp.cuts)
    if (c.first == task_units) {
      cp.d_tasks_own = c.second.first;
      cp.ntasks = c.second.second;
      return ASUKA_OK;
    }
  const std::vector<RowTask> tk =
      make_tasks(cp, ctx->h_units.data() + cp.unit_start, ctx->h_col, ctx->pp_count, task_units);
  RowTask* d = nullptr;
  HIP_CHECK(hipMalloc(&d, sizeof(RowTask) * std::max<size_t>(1, tk.size())));
  ctx->owned.push_back(d);
  HIP_CHECK(hipMemcpyAsync(d, tk.data(), sizeof(RowTask) * tk.size(), hipMemcpyHostToDevice, s));
  HIP_CHECK(hipStreamSynchronize(s));
  cp.d_tasks_own = d;
  cp.ntasks = (long long)tk.size();
  cp.cuts.push_back({task_units, {d, cp.ntasks}});
  return ASUKA_OK;
}

// Applies (wavefronts per workgroup, units per task, LDS accumulators) to class pair i.
static int apply_shape(asuka_jk_ctx* ctx, int i, int wpg, int task_units, bool acc, hipStream_t s) {
  ClassPairHost& cp = ctx->cps[i];
  int cC = 0, cD = 0;
  const bool fits = row_acc_fits(ctx, cp.la, cp.lb, cp.lc, cp.ld, &cC, &cD);
  cp.acc_colsC = (acc && fits) ? cC : 0;
  cp.acc_colsD = (acc && fits) ? cD : 0;
  cp.wpg = wpg;
  return recut_tasks(ctx, i, task_units, s);
}

extern "C" {

int asuka_jk_autotune(asuka_jk_ctx* ctx, const double* D, int flags, int reps, void* stream) {
  if (!ctx || !D) return fail(ASUKA_E_INVALID, "ctx and D must be non-null");
  if (reject_batch(ctx)) return ASUKA_E_INVALID;
  if (reps < 1) return fail(ASUKA_E_INVALID, "reps must be >= 1");
  hipStream_t s = (hipStream_t)stream;
  const long long nn = (long long)ctx->nao * ctx->nao;
  double* tjk = nullptr;  // J and K of the timed launches (discarded)
  HIP_CHECK(hipMalloc(&tjk, sizeof(double) * 2 * nn));
  struct Free {
    double* p;
    ~Free() { (void)hipFree(p); }
  } free_tjk{tjk};
  hipEvent_t e0 = nullptr, e1 = nullptr;
  HIP_CHECK(hipEventCreate(&e0));
  HIP_CHECK(hipEventCreate(&e1));
  struct FreeEv {
    hipEvent_t a, b;
    ~FreeEv() {
      (void)hipEventDestroy(a);
      (void)hipEventDestroy(b);
    }
  } free_ev{e0, e1};
  hipLaunchKernelGGL(k_dsp, dim3(grid_for(ctx->nsp, 128)), dim3(128), 0, s, D, nullptr, ctx->dev, ctx->d_Dsp,
                     ctx->d_kd, ctx->d_pos_of_sp);
  HIP_CHECK(hipGetLastError());
  EriOut o{};
  o.mode = MODE_JK;
  o.D = D;
  o.J = tjk;
  o.K = tjk + nn;
  o.eps_density = ctx->eps_density;
  o.eps_prim = ctx->eps_schwarz > 0.0 ? ctx->eps_prim : 0.0;
  o.mixed = (flags & ASUKA_JK_MIXED) ? 1 : 0;
  o.tau_mixed = ctx->tau_mixed;
  // median of reps timed launches (after one untimed) of class pair i as it is now set up
  auto time_class = [&](int i, double* ms) -> int {
    std::vector<float> t;
    for (int r = 0; r <= reps; ++r) {
      ClassWork w = work_of(ctx, ctx->cps[i], i);
      hipLaunchKernelGGL(k_zero_u32, dim3(1), dim3(64), 0, s, w.queue, 1LL);
      hipLaunchKernelGGL(k_zero_u32, dim3(1), dim3(64), 0, s, w.queue32, 1LL);
      HIP_CHECK(hipEventRecord(e0, s));
      bool handled = false;
      HIP_CHECK(launch_eri_fast(ctx->dev, w, o, s, &handled));
      if (!handled) HIP_CHECK(launch_eri_coop(ctx->dev, w, o, s, &handled));
      if (!handled) HIP_CHECK(launch_eri_generic(ctx->dev, w, o, s));
      HIP_CHECK(hipEventRecord(e1, s));
      HIP_CHECK(hipEventSynchronize(e1));
      float f = 0.0f;
      HIP_CHECK(hipEventElapsedTime(&f, e0, e1));
      if (r > 0) t.push_back(f);
    }
    std::sort(t.begin(), t.end());
    *ms = (double)t[t.size() / 2];
    return ASUKA_OK;
  };
  int rc = ASUKA_OK;
  for (size_t ii = 0; ii < ctx->cps.size(); ++ii) {
    const int i = (int)ii;
    ClassPairHost& cp = ctx->cps[i];
    double best = 0.0;
    if ((rc = time_class(i, &best))) return rc;
    if (eri_fast_has_plan(cp.la, cp.lb, cp.lc, cp.ld)) {
      // candidate shapes: 4 or 8 wavefronts per workgroup, 16..128 wave units
      // per row task, LDS row accumulators on (where they fit) or off
      int cC = 0, cD = 0;
      const bool fits = row_acc_fits(ctx, cp.la, cp.lb, cp.lc, cp.ld, &cC, &cD);
      int bw = cp.wpg, bt = cp.task_units;
      bool ba = cp.acc_colsC > 0;
      for (int acc = fits ? 1 : 0; acc >= 0; --acc)
        for (int wpg : {4, 8})
          for (int tu : {16, 32, 64, 128}) {
            if ((rc = apply_shape(ctx, i, wpg, tu, acc != 0, s))) return rc;
            double ms = 0.0;
            if ((rc = time_class(i, &ms))) return rc;
            if (ms < best) {
              best = ms;
              bw = wpg;
              bt = tu;
              ba = acc != 0;
            }
          }
      if ((rc = apply_shape(ctx, i, bw, bt, ba, s))) return rc;
    }
    cp.tuned_ms = std::max(best, 1e-6);
  }
  assign_streams(ctx);
  return route_task_layout(ctx);  // the cuts may have changed
}

int asuka_jk_tuning_get(const asuka_jk_ctx* ctx, int32_t* rows, int cap) {
  if (!ctx || (cap > 0 && !rows)) return fail(ASUKA_E_INVALID, "ctx and rows must be non-null");
  const int n = std::min<int>(cap, (int)ctx->cps.size());
  for (int i = 0; i < n; ++i) {
    const ClassPairHost& cp = ctx->cps[i];
    int32_t* r = rows + 8 * i;
    r[0] = cp.la;
    r[1] = cp.lb;
    r[2] = cp.lc;
    r[3] = cp.ld;
    r[4] = cp.wpg;
    r[5] = cp.task_units;
    r[6] = cp.acc_colsC > 0 ? 1 : 0;
    r[7] = (int32_t)std::llround(cp.tuned_ms * 1000.0);
  }
  return (int)ctx->cps.size();
}

int asuka_jk_tuning_set(asuka_jk_ctx* ctx, const int32_t* rows, int n, void* stream) {
  if (!ctx || (n > 0 && !rows)) return fail(ASUKA_E_INVALID, "ctx and rows must be non-null");
  hipStream_t s = (hipStream_t)stream;
  for (int k = 0; k < n; ++k) {
    const int32_t* r = rows + 8 * k;
    if (r[4] != 4 && r[4] != 8) return fail(ASUKA_E_INVALID, "wavefronts per workgroup must be 4 or 8");
    if (r[5] < 1 || r[5] > 1024) return fail(ASUKA_E_INVALID, "units per task must be in [1, 1024]");
  }
  for (int k = 0; k < n; ++k) {
    const int32_t* r = rows + 8 * k;
    for (size_t i = 0; i < ctx->cps.size(); ++i) {
      ClassPairHost& cp = ctx->cps[i];
      if (cp.la != r[0] || cp.lb != r[1] || cp.lc != r[2] || cp.ld != r[3]) continue;
      const int rc = apply_shape(ctx, (int)i, r[4], r[5], r[6] != 0, s);
      if (rc) return rc;
      cp.tuned_ms = r[7] > 0 ? r[7] / 1000.0 : 0.0;
    }
  }
  assign_streams(ctx);
  return route_task_layout(ctx);  // the cuts may have changed
}

int asuka_jk_last_timing(asuka_jk_ctx* ctx, double* ms, int64_t* nlaunch) {
  if (!ctx || !ms) return fail(ASUKA_E_INVALID, "ctx and ms must be non-null");
  if (!ctx->ev_start || !ctx->ev_stop) return fail(ASUKA_E_INVALID, "no timed build recorded (use ASUKA_JK_TIME)");
  HIP_CHECK(hipEventSynchronize(ctx->ev_stop));
  float f = 0.0f;
  HIP_CHECK(hipEventElapsedTime(&f, ctx->ev_start, ctx->ev_stop));
  *ms = (double)f;
  if (nlaunch) *nlaunch = ctx->last_launches;
  return ASUKA_OK;
}

int asuka_jk_last_counts(asuka_jk_ctx* ctx, int64_t* rows, int cap, void* stream) {
  if (!ctx || (cap > 0 && !rows)) return fail(ASUKA_E_INVALID, "ctx and rows must be non-null");
  const int n = std::min<int>(cap, (int)ctx->cps.size());
  if (n <= 0) return 0;
  std::vector<unsigned long long> h(kCountStride * (size_t)n);
  HIP_CHECK(hipMemcpyAsync(h.data(), ctx->d_counters, sizeof(unsigned long long) * h.size(), hipMemcpyDeviceToHost,
                           (hipStream_t)stream));
  HIP_CHECK(hipStreamSynchronize((hipStream_t)stream));
  for (size_t i = 0; i < h.size(); ++i) rows[i] = (int64_t)h[i];
  return n;
}

int asuka_jk_stats(asuka_jk_ctx* ctx, asuka_jk_stats_out* out, void* stream) {
  if (!ctx || !out) return fail(ASUKA_E_INVALID, "ctx and out must be non-null");
  *out = asuka_jk_stats_out{};
  const int n = (int)ctx->cps.size();
  out->class_pairs = n;
  std::vector<unsigned long long> h(kCountStride * (size_t)std::max(1, n), 0ull);
  if (n > 0) {
    HIP_CHECK(hipMemcpyAsync(h.data(), ctx->d_counters, sizeof(unsigned long long) * kCountStride * n,
                             hipMemcpyDeviceToHost, (hipStream_t)stream));
    HIP_CHECK(hipStreamSynchronize((hipStream_t)stream));
  }
  for (int i = 0; i < n; ++i) {
    const ClassPairHost& cp = ctx->cps[i];
    const unsigned long long* c = h.data() + (size_t)kCountStride * i;
    out->schwarz_quartets += cp.nquartets;
    // EFM-v1: per primitive quartet 24 + n (24 + 15 (nmax+1)(mmax+1) + 6 N1 + 4 ncomp);
    // per quartet 2 ncomp (jt + kt) for the digestion
    const int nr = (cp.la + cp.lb + cp.lc + cp.ld) / 2 + 1;
    const double ncomp = (double)ncart_c(cp.la) * ncart_c(cp.lb) * ncart_c(cp.lc) * ncart_c(cp.ld);
    const double n1 = (double)(cp.la + 1) * (cp.lb + 1) * (cp.lc + 1) * (cp.ld + 1);
    const double per_prim = 24.0 + nr * (24.0 + 15.0 * (cp.la + cp.lb + 1) * (cp.lc + cp.ld + 1) + 6.0 * n1 + 4.0 * ncomp);
    for (int bin = 0; bin < 8; ++bin) {
      const int ab = (bin >> 2) & 1, cd = (bin >> 1) & 1, pq = bin & 1;
      const int jt = 1 + pq, kt = 1 + cd + ab + (ab & cd);
      out->quartets += (int64_t)c[bin];
      out->flops_efm += 2.0 * ncomp * (jt + kt) * (double)c[bin];
    }
    out->prim_quartets += (int64_t)c[8];
    out->quartets_fp32 += (int64_t)c[9];
    out->prim_quartets_fp32 += (int64_t)c[10];
    out->flops_efm += per_prim * (double)c[8];
    out->flops_efm_fp32 += per_prim * (double)c[10];
  }
  if (ctx->ev_start && ctx->ev_stop && hipEventQuery(ctx->ev_stop) != hipErrorNotReady) {
    float f = 0.0f;
    if (hipEventElapsedTime(&f, ctx->ev_start, ctx->ev_stop) == hipSuccess) out->last_ms = (double)f;
  }
  return ASUKA_OK;
}

int asuka_jk_eri_tiles(asuka_jk_ctx* ctx, int64_t nq, const int32_t* spab, const int32_t* spcd, const int64_t* out_off,
                       double* out, int flags, void* stream) {
  if (!ctx || nq < 0 || (nq > 0 && (!spab || !spcd || !out_off || !out)))
    return fail(ASUKA_E_INVALID, "tile arguments must be non-null");
  (void)flags;
  if (nq == 0) return ASUKA_OK;
  ClassWork w{};
  w.la = w.lb = w.lc = w.ld = ctx->max_l;
  w.nwork = nq;
  w.q_ab = spab;
  w.q_cd = spcd;
  w.q_out_off = (const long long*)out_off;
  EriOut o{};
  o.mode = MODE_TILE;
  o.tile = out;
  HIP_CHECK(launch_eri_generic(ctx->dev, w, o, (hipStream_t)stream));
  return ASUKA_OK;
}

int asuka_int1e(asuka_jk_ctx* ctx, double* S, double* T, double* V, void* stream) {
  if (!ctx) return fail(ASUKA_E_INVALID, "ctx must be non-null");
  if (reject_batch(ctx)) return ASUKA_E_INVALID;
  if (V && ctx->natm <= 0) return fail(ASUKA_E_INVALID, "V needs atoms (natm > 0)");
  HIP_CHECK(launch_int1e(ctx, S, T, V, (hipStream_t)stream));
  return ASUKA_OK;
}

int asuka_jk_grad(asuka_jk_ctx* ctx, const double* D, const double* WJ, const double* WK, double* grad,
                  void* stream) {
  if (!ctx || !D || !grad) return fail(ASUKA_E_INVALID, "ctx, D and grad must be non-null");
  if (reject_batch(ctx)) return ASUKA_E_INVALID;
  if (ctx->natm <= 0) return fail(ASUKA_E_INVALID, "the gradient needs atoms (natm > 0)");
  hipStream_t s = (hipStream_t)stream;
  hipLaunchKernelGGL(k_fill, dim3(1), dim3(256), 0, s, grad, (long long)3 * ctx->natm, 0.0);
  HIP_CHECK(hipGetLastError());
  if (!WJ && !WK) return ASUKA_OK;
  {  // work-queue heads of the lane-per-quartet kernels start at zero
    const long long nq = (long long)std::max<size_t>(1, ctx->cps.size());
    hipLaunchKernelGGL(k_zero_u32, dim3((unsigned)((nq + 255) / 256)), dim3(256), 0, s, ctx->d_queue, nq);
    HIP_CHECK(hipGetLastError());
  }
  int rc = fork_streams(ctx, s);  // class pairs spread over the context's streams, as in the build
  if (rc) return rc;
  for (int i : ctx->launch_order) {
    hipStream_t ls = stream_of(ctx, ctx->cps[i], s);
    ClassWork w = work_of(ctx, ctx->cps[i], i);
    bool handled = false;
    HIP_CHECK(launch_eri_deriv_fast(ctx->dev, w, D, WJ, WK, ctx->d_shell_atom, ctx->natm, grad, ls, &handled));
    if (!handled) HIP_CHECK(launch_eri_deriv(ctx->dev, w, D, WJ, WK, ctx->d_shell_atom, grad, ls));
  }
  return join_streams(ctx, s);
}

int asuka_int1e_grad(asuka_jk_ctx* ctx, const double* Dh, const double* Ws, double* grad, void* stream) {
  if (!ctx || !grad) return fail(ASUKA_E_INVALID, "ctx and grad must be non-null");
  if (reject_batch(ctx)) return ASUKA_E_INVALID;
  if (ctx->natm <= 0) return fail(ASUKA_E_INVALID, "the gradient needs atoms (natm > 0)");
  hipStream_t s = (hipStream_t)stream;
  hipLaunchKernelGGL(k_fill, dim3(1), dim3(256), 0, s, grad, (long long)3 * ctx->natm, 0.0);
  HIP_CHECK(hipGetLastError());
  if (!Dh && !Ws) return ASUKA_OK;
  HIP_CHECK(launch_int1e_grad(ctx->dev, ctx->d_prim_start, ctx->d_prim_num, ctx->d_prim_exp, ctx->d_prim_coef,
                              ctx->natm, ctx->d_atom_xyz, ctx->d_atom_Z, ctx->d_shell_atom, Dh, Ws, grad, s));
  return ASUKA_OK;
}

int asuka_rys_roots(int n, double T, double* x, double* w) {
  if (n < 1 || n > RYS_NMAX) return fail(ASUKA_E_INVALID, "n must be in [1, 7]");
  if (!(T >= 0.0)) return fail(ASUKA_E_INVALID, "T must be >= 0");
  if (!x || !w) return fail(ASUKA_E_INVALID, "x and w must be non-null");
  for (int u = 0; u < n; ++u) rys_root_host(n, u, T, x[u], w[u]);
  return ASUKA_OK;
}

int asuka_rys_roots_device(int n, int64_t npts, const double* T, double* x, double* w, void* stream) {
  if (n < 1 || n > RYS_NMAX) return fail(ASUKA_E_INVALID, "n must be in [1, 7]");
  if (npts < 0 || (npts > 0 && (!T || !x || !w))) return fail(ASUKA_E_INVALID, "arrays must be non-null");
  if (npts == 0) return ASUKA_OK;
  hipLaunchKernelGGL(k_rys_points, dim3((unsigned)((npts + 127) / 128)), dim3(128), 0, (hipStream_t)stream, n,
                     (long long)npts, T, x, w);
  HIP_CHECK(hipGetLastError());
  return ASUKA_OK;
}

}  // extern "C"

// ---------------------------------------------------------------------------
// One-electron integrals, one thread per shell pair.
// S and T: 1-D overlap recursion and the kinetic-from-overlap relation
// (asuka/integrals/int1e_cart.py:77-115, :758-787). V: Rys quadrature with a
// point-charge ket (the q -> infinity limit of the ERI recursion):
// G[a+1] = (PA - x PC) G[a] + a (1-x)/(2p) G[a-1], HRR on the bra, weight
// -Z 2 pi / p exp(-mu AB^2) sum_u w_u Ix Iy Iz, T = p |PC|^2; equal to the
// reference's McMurchie-Davidson form (int1e_cart.py:1031-1190).
// ---------------------------------------------------------------------------
namespace {

constexpr int L1 = ASUKA_LMAX;
constexpr int LT = L1 + 3;  // overlap table needs lb + 2

__device__ void overlap_1d(int la, int lb, double a, double b, double A, double B, double S[L1 + 1][LT]) {
  const double p = a + b, ip = 1.0 / p, mu = a * b * ip;
  const double Px = (a * A + b * B) * ip, PA = Px - A, PB = Px - B, AB = A - B;
  for (int i = 0; i <= L1; ++i)
    for (int j = 0; j < LT; ++j) S[i][j] = 0.0;
  S[0][0] = sqrt(3.14159265358979323846 * ip) * exp(-mu * AB * AB);
  const double i2p = 0.5 * ip;
  for (int j = 0; j < lb; ++j) S[0][j + 1] = PB * S[0][j] + (j > 0 ? j * i2p * S[0][j - 1] : 0.0);
  for (int i = 0; i < la; ++i) {
    S[i + 1][0] = PA * S[i][0] + (i > 0 ? i * i2p * S[i - 1][0] : 0.0);
    for (int j = 0; j < lb; ++j)
      S[i + 1][j + 1] = PA * S[i][j + 1] + (i > 0 ? i * i2p * S[i - 1][j + 1] : 0.0) + (j + 1) * i2p * S[i][j];
  }
}

__device__ double kin1d(double S[L1 + 1][LT], int i, int j, double b) {
  double v = b * (2.0 * j + 1.0) * S[i][j] - 2.0 * b * b * S[i][j + 2];
  if (j >= 2) v -= 0.5 * (double)(j * (j - 1)) * S[i][j - 2];
  return v;
}

// 1-D nuclear-attraction factor for one root: I(i, j), i <= la, j <= lb.
__device__ double nuc1d(int i, int j, double PA, double PC, double AB, double x, double i2p) {
  double G[2 * L1 + 1];
  const double c0 = PA - x * PC, b1 = (1.0 - x) * i2p;
  G[0] = 1.0;
  const int n = i + j;
  if (n > 0) G[1] = c0;
  for (int a = 1; a < n; ++a) G[a + 1] = c0 * G[a] + a * b1 * G[a - 1];
  // (i, j) = sum_k C(j,k) AB^(j-k) G[i+k]
  double out = 0.0, binom = 1.0;
  for (int k = 0; k <= j; ++k) {
    double pw = 1.0;
    for (int e = 0; e < j - k; ++e) pw *= AB;
    out += binom * pw * G[i + k];
    binom = binom * (double)(j - k) / (double)(k + 1);
  }
  return out;
}

__global__ void k_int1e(const DevBasis b, const int* pstart, const int* pnum, const double* pexp, const double* pcoef,
                        int natm, const double* axyz, const double* aZ, double* S, double* T, double* V) {
  const int sp = blockIdx.x * blockDim.x + threadIdx.x;
  if (sp >= b.nsp) return;
  const int A = b.sp_A[sp], B = b.sp_B[sp];
  const int la = b.shell_l[A], lb = b.shell_l[B];
  const int nA = ncart_c(la), nB = ncart_c(lb);
  int ax[10], ay[10], az[10], bx[10], by[10], bz[10];
  {
    int i = 0;
    for (int x = la; x >= 0; --x)
      for (int y = la - x; y >= 0; --y) {
        ax[i] = x;
        ay[i] = y;
        az[i] = la - x - y;
        ++i;
      }
    i = 0;
    for (int x = lb; x >= 0; --x)
      for (int y = lb - x; y >= 0; --y) {
        bx[i] = x;
        by[i] = y;
        bz[i] = lb - x - y;
        ++i;
      }
  }
  double s[100], t[100], v[100];
  for (int i = 0; i < nA * nB; ++i) s[i] = t[i] = v[i] = 0.0;
  const double* cA = b.shell_xyz + 3 * A;
  const double* cB = b.shell_xyz + 3 * B;
  const double ABx = cA[0] - cB[0], ABy = cA[1] - cB[1], ABz = cA[2] - cB[2];
  const double AB2 = ABx * ABx + ABy * ABy + ABz * ABz;
  const int nr = (la + lb) / 2 + 1;
  for (int ia_p = 0; ia_p < pnum[A]; ++ia_p)
    for (int ib_p = 0; ib_p < pnum[B]; ++ib_p) {
      const double a = pexp[pstart[A] + ia_p], bb = pexp[pstart[B] + ib_p];
      const double c = pcoef[pstart[A] + ia_p] * pcoef[pstart[B] + ib_p];
      if (c == 0.0) continue;
      double Sx[L1 + 1][LT], Sy[L1 + 1][LT], Sz[L1 + 1][LT];
      overlap_1d(la, lb + 2, a, bb, cA[0], cB[0], Sx);
      overlap_1d(la, lb + 2, a, bb, cA[1], cB[1], Sy);
      overlap_1d(la, lb + 2, a, bb, cA[2], cB[2], Sz);
      for (int i = 0; i < nA; ++i)
        for (int j = 0; j < nB; ++j) {
          const double sx = Sx[ax[i]][bx[j]], sy = Sy[ay[i]][by[j]], sz = Sz[az[i]][bz[j]];
          s[i * nB + j] += c * sx * sy * sz;
          t[i * nB + j] += c * (kin1d(Sx, ax[i], bx[j], bb) * sy * sz + sx * kin1d(Sy, ay[i], by[j], bb) * sz +
                                sx * sy * kin1d(Sz, az[i], bz[j], bb));
        }
      if (!V) continue;
      const double p = a + bb, ip = 1.0 / p;
      const double Px = (a * cA[0] + bb * cB[0]) * ip, Py = (a * cA[1] + bb * cB[1]) * ip,
                   Pz = (a * cA[2] + bb * cB[2]) * ip;
      const double Kab = exp(-a * bb * ip * AB2);
      for (int ic = 0; ic < natm; ++ic) {
        const double Z = aZ[ic];
        if (Z == 0.0) continue;
        const double PCx = Px - axyz[3 * ic], PCy = Py - axyz[3 * ic + 1], PCz = Pz - axyz[3 * ic + 2];
        const double Targ = p * (PCx * PCx + PCy * PCy + PCz * PCz);
        const double pref = -Z * 2.0 * 3.14159265358979323846 * ip * Kab * c;
        for (int u = 0; u < nr; ++u) {
          double xr, wr;
          rys_root_dev(nr, u, Targ, xr, wr);
          for (int i = 0; i < nA; ++i)
            for (int j = 0; j < nB; ++j) {
              const double Ix = nuc1d(ax[i], bx[j], Px - cA[0], PCx, ABx, xr, 0.5 * ip);
              const double Iy = nuc1d(ay[i], by[j], Py - cA[1], PCy, ABy, xr, 0.5 * ip);
              const double Iz = nuc1d(az[i], bz[j], Pz - cA[2], PCz, ABz, xr, 0.5 * ip);
              v[i * nB + j] += pref * wr * Ix * Iy * Iz;
            }
        }
      }
    }
  const long long N = b.nao;
  const int a0 = b.shell_ao[A], b0 = b.shell_ao[B];
  for (int i = 0; i < nA; ++i)
    for (int j = 0; j < nB; ++j) {
      const long long r = a0 + i, cc = b0 + j;
      if (S) S[r * N + cc] = S[cc * N + r] = s[i * nB + j];
      if (T) T[r * N + cc] = T[cc * N + r] = t[i * nB + j];
      if (V) V[r * N + cc] = V[cc * N + r] = v[i * nB + j];
    }
}

}  // namespace

hipError_t launch_int1e(const asuka_jk_ctx* ctx, double* S, double* T, double* V, hipStream_t s) {
  hipLaunchKernelGGL(k_int1e, dim3((ctx->nsp + 63) / 64), dim3(64), 0, s, ctx->dev, ctx->d_prim_start,
                     ctx->d_prim_num, ctx->d_prim_exp, ctx->d_prim_coef, ctx->natm, ctx->d_atom_xyz, ctx->d_atom_Z, S,
                     T, V);
  return hipGetLastError();
}
