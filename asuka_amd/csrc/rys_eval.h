// Rys roots and weights from the generated interpolation table.
//
// The table data are included twice, once for device code (namespace rys_dev,
// __device__ constant data) and once for host code (namespace rys_host), so the
// same evaluator serves kernels and the host-side test hook.
//
// Roots are x = t^2 in [0, 1]; sum_u w_u x_u^k = F_k(T) for k < 2n.
#pragma once

#include <hip/hip_runtime.h>

namespace rys_dev {
#define RYS_TABLE_QUAL static const __device__
#define RYS_COEF_T double
#include "rys_table.h"
#undef RYS_COEF_T
#undef RYS_TABLE_QUAL
}  // namespace rys_dev

// The FP32 path's own table (gen_rys_table_f32.py): degree-RYS32_DEG
// polynomials on the same intervals up to its own, lower asymptotic switch.
namespace rys_dev_f {
#define RYS_TABLE_QUAL static const __device__
#include "rys_table_f32.h"
#undef RYS_TABLE_QUAL
}  // namespace rys_dev_f

namespace rys_host {
#define RYS_TABLE_QUAL static const
#define RYS_COEF_T double
#include "rys_table.h"
#undef RYS_COEF_T
#undef RYS_TABLE_QUAL
}  // namespace rys_host

// Evaluate root u (0-based, ascending) of the n-point rule at T.
// Host and device variants differ only in which copy of the table they read.
#define ASUKA_DEFINE_RYS_ONE(NAME, QUAL, NS)                                                  \
  QUAL inline void NAME(int n, int u, double T, double& x, double& w) {                     \
    if (T >= NS::rys_tmax[n]) {                                                             \
      const double it = 1.0 / T;                                                            \
      const int o = NS::rys_herm_off[n] + u;                                                \
      x = NS::rys_herm_x[o] * it;                                                           \
      w = NS::rys_herm_w[o] * sqrt(it);                                                     \
      return;                                                                               \
    }                                                                                       \
    int i = (int)(T * RYS_INV_H);                                                           \
    if (i >= NS::rys_nint[n]) i = NS::rys_nint[n] - 1;                                     \
    const double s = (T - (double)i * RYS_H) * (2.0 * RYS_INV_H) - 1.0;                     \
    const double* c = NS::rys_coef + NS::rys_coef_off[n] + (i * 2 * n + u) * (RYS_DEG + 1); \
    const double* cw = c + n * (RYS_DEG + 1);                                               \
    double ax = c[RYS_DEG], aw = cw[RYS_DEG];                                               \
    _Pragma("unroll") for (int d = RYS_DEG - 1; d >= 0; --d) {                              \
      ax = fma(ax, s, c[d]);                                                                \
      aw = fma(aw, s, cw[d]);                                                               \
    }                                                                                       \
    x = ax;                                                                                 \
    w = aw;                                                                                 \
  }

ASUKA_DEFINE_RYS_ONE(rys_root_dev, __device__, rys_dev)
ASUKA_DEFINE_RYS_ONE(rys_root_host, , rys_host)

// All n roots at once (compile-time n), device.
template <int N>
__device__ inline void rys_roots_dev(double T, double* x, double* w) {
  if (T >= rys_dev::rys_tmax[N]) {
    const double it = 1.0 / T;
    const double sit = sqrt(it);
#pragma unroll
    for (int u = 0; u < N; ++u) {
      x[u] = rys_dev::rys_herm_x[rys_dev::rys_herm_off[N] + u] * it;
      w[u] = rys_dev::rys_herm_w[rys_dev::rys_herm_off[N] + u] * sit;
    }
    return;
  }
  int i = (int)(T * RYS_INV_H);
  if (i >= rys_dev::rys_nint[N]) i = rys_dev::rys_nint[N] - 1;
  const double s = (T - (double)i * RYS_H) * (2.0 * RYS_INV_H) - 1.0;
  const double* c = rys_dev::rys_coef + rys_dev::rys_coef_off[N] + i * 2 * N * (RYS_DEG + 1);
#pragma unroll
  for (int f = 0; f < 2 * N; ++f) {
    const double* cf = c + f * (RYS_DEG + 1);
    double a = cf[RYS_DEG];
#pragma unroll
    for (int d = RYS_DEG - 1; d >= 0; --d) a = fma(a, s, cf[d]);
    if (f < N)
      x[f] = a;
    else
      w[f - N] = a;
  }
}

// Compile-time copies of the table metadata.
namespace rys_c {
constexpr int nint[] = {RYS_NINT_LIST};
constexpr double tmax[] = {RYS_TMAX_LIST};
constexpr int coef_off[] = {RYS_COEF_OFF_LIST};
// coefficients of the n-root block
constexpr int block_size(int n) { return nint[n] * 2 * n * (RYS_DEG + 1); }
// coefficients of one interval of the n-root block
constexpr int interval_size(int n) { return 2 * n * (RYS_DEG + 1); }
// Interval stride (elements of size esize) of an LDS copy of the n-root block:
// one interval's coefficients padded to an odd number of 16-byte slots, so
// lanes whose T falls in different intervals read different LDS banks with
// their 16-byte reads (same-interval lanes share addresses and broadcast).
constexpr int lds_stride(int n, int esize) {
  const int per16 = 16 / esize;
  int s = (interval_size(n) + per16 - 1) / per16 * per16;
  if ((s / per16) % 2 == 0) s += per16;
  return s;
}
// LDS bytes of the padded FP64 copy of the n-root block (the larger of a
// kernel's two possible copies)
constexpr size_t lds_table_bytes(int n) { return (size_t)nint[n] * (size_t)lds_stride(n, 8) * 8; }

// The table a kernel of arithmetic CT reads: the FP64 one, or for float the
// FP32 table (degree RYS32_DEG, its own asymptotic switch and interval count).
constexpr int nint32[] = {RYS32_NINT_LIST};
constexpr double tmax32[] = {RYS32_TMAX_LIST};
constexpr int coef_off32[] = {RYS32_COEF_OFF_LIST};
template <typename CT>
constexpr int deg_of() { return sizeof(CT) == 4 ? RYS32_DEG : RYS_DEG; }
template <typename CT>
constexpr int nint_of(int n) { return sizeof(CT) == 4 ? nint32[n] : nint[n]; }
template <typename CT>
constexpr double tmax_of(int n) { return sizeof(CT) == 4 ? tmax32[n] : tmax[n]; }
template <typename CT>
constexpr int coef_off_of(int n) { return sizeof(CT) == 4 ? coef_off32[n] : coef_off[n]; }
template <typename CT>
constexpr int interval_size_of(int n) { return 2 * n * (deg_of<CT>() + 1); }
template <typename CT>
constexpr int block_size_of(int n) { return nint_of<CT>(n) * interval_size_of<CT>(n); }
// interval stride of a padded LDS copy, as lds_stride
template <typename CT>
constexpr int lds_stride_of(int n) {
  constexpr int per16 = 16 / (int)sizeof(CT);
  int s = (interval_size_of<CT>(n) + per16 - 1) / per16 * per16;
  if ((s / per16) % 2 == 0) s += per16;
  return s;
}
static_assert(RYS_H == 2.0, "the block evaluators assume intervals of width 2");
}  // namespace rys_c

// Reciprocal and reciprocal square root in FP64: the hardware estimate
// (v_rcp_f64 / v_rsq_f64) refined by two Newton-Raphson steps, each of which
// squares the relative error, to within an ulp or two of the correctly
// rounded result. Valid for normal positive arguments far from the
// overflow/underflow limits (pair exponents and Boys arguments are).
__device__ __forceinline__ double rcp_nr(double x) {
  double r = __builtin_amdgcn_rcp(x);
  double e = fma(-x, r, 1.0);
  r = fma(r, e, r);
  e = fma(-x, r, 1.0);
  return fma(r, e, r);
}
__device__ __forceinline__ double rsq_nr(double x) {
  double y = __builtin_amdgcn_rsq(x);
  const double h = 0.5 * x;
  double t = fma(-h * y, y, 0.5);  // y <- y (3 - x y^2) / 2
  y = fma(y, t, y);
  t = fma(-h * y, y, 0.5);
  return fma(y, t, y);
}
__device__ __forceinline__ float rcp_nr(float x) { return __builtin_amdgcn_rcpf(x); }
__device__ __forceinline__ float rsq_nr(float x) { return __builtin_amdgcn_rsqf(x); }

// All N roots and weights at T from the N-root coefficient block `blk` of
// the CT table (rys_coef + coef_off[N] for double, rys32_coef + coef_off32[N]
// for float, or a copy of it in LDS), in the arithmetic of CT. At T >=
// tmax_of<CT>(N) the Gauss-Hermite limit, with 1/T and 1/sqrt(T) in CT.
// STRIDE is the interval stride of blk: interval_size_of<CT>(N) for the global
// table, lds_stride_of<CT>(N) for a padded LDS copy.
template <int N, typename CT, int STRIDE = rys_c::interval_size_of<CT>(N)>
__device__ __forceinline__ void rys_roots_blk(CT T, const CT* __restrict__ blk, CT* x, CT* w) {
  constexpr int NI = rys_c::nint_of<CT>(N);
  constexpr int DEG = rys_c::deg_of<CT>();
  if (T >= CT(rys_c::tmax_of<CT>(N))) {
    const CT it = rcp_nr(T);
    const CT sit = rsq_nr(T);
#pragma unroll
    for (int u = 0; u < N; ++u) {
      x[u] = CT(rys_dev::rys_herm_x[rys_dev::rys_herm_off[N] + u]) * it;
      w[u] = CT(rys_dev::rys_herm_w[rys_dev::rys_herm_off[N] + u]) * sit;
    }
    return;
  }
  int i = (int)(T * CT(0.5));
  i = i < NI - 1 ? i : NI - 1;
  const CT s = (T - CT(2 * i)) - CT(1);
  const CT* c = blk + i * STRIDE;
#pragma unroll
  for (int f = 0; f < 2 * N; ++f) {
    const CT* cf = c + f * (DEG + 1);
    CT a = cf[DEG];
#pragma unroll
    for (int d = DEG - 1; d >= 0; --d) a = fma(a, s, cf[d]);
    if (f < N)
      x[f] = a;
    else
      w[f - N] = a;
  }
}
