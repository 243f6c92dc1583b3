// Device side of the graph-captured SCF iteration: the DIIS extrapolation
// coefficients (asuka/hf/df_scf.py:722-835) solved on the GPU, so the SCF
// iteration needs no host round trip.
//
// The driver keeps up to kDiisSlots Fock matrices and commutator errors in a
// ring of slots and their Gram matrix G[i][j] = <e_i, e_j> on the device;
// this kernel forms the reference's bordered system
//   [ G/g  -1 ] [c]   [ 0]
//   [ -1    0 ] [l] = [-1],   g = max |G|,
// solves it by Gaussian elimination with partial pivoting, retrying with
// lambda I added to G/g for lambda in (0, 1e-14, 1e-12, 1e-10, 1e-8, 1e-6)
// until the solve has no zero pivot and a finite result, and renormalises c
// to unit sum. Where the reference would fall back (fewer than 2 vectors, a
// non-finite G, every solve failing, |sum c| < 1e-15), the weights select the
// newest Fock matrix alone; the reference's least-squares retry after every
// solve failed is not taken (it cannot be reached with a finite G, whose
// regularised systems are nonsingular).
#include <hip/hip_runtime.h>

#include <cmath>
#include <string>

#include "../../include/asuka_jk.h"

namespace {

constexpr int kDiisSlots = 8;

// Solves the (m x m) system A x = b in place (A row-major, partial pivoting);
// false on an exactly zero pivot (numpy/LAPACK getrf info > 0) or a non-finite x.
__device__ bool solve_small(double* A, double* b, int m) {
  for (int k = 0; k < m; ++k) {
    int p = k;
    double pv = fabs(A[k * m + k]);
    for (int i = k + 1; i < m; ++i)
      if (fabs(A[i * m + k]) > pv) {
        pv = fabs(A[i * m + k]);
        p = i;
      }
    if (pv == 0.0) return false;
    if (p != k) {
      for (int j = 0; j < m; ++j) {
        const double t = A[k * m + j];
        A[k * m + j] = A[p * m + j];
        A[p * m + j] = t;
      }
      const double t = b[k];
      b[k] = b[p];
      b[p] = t;
    }
    for (int i = k + 1; i < m; ++i) {
      const double f = A[i * m + k] / A[k * m + k];
      for (int j = k; j < m; ++j) A[i * m + j] -= f * A[k * m + j];
      b[i] -= f * b[k];
    }
  }
  for (int k = m - 1; k >= 0; --k) {
    double s = b[k];
    for (int j = k + 1; j < m; ++j) s -= A[k * m + j] * b[j];
    b[k] = s / A[k * m + k];
    if (!isfinite(b[k])) return false;
  }
  return true;
}

// One thread: G (kDiisSlots x kDiisSlots, row stride kDiisSlots) over a ring
// of nslot <= kDiisSlots slots, count = number of vectors pushed so far (slots
// 0..n-1 in use, n = min(count, nslot); the newest in slot (count - 1) % nslot).
// coeff (kDiisSlots): extrapolation weights per slot; status: 1 extrapolated,
// 0 newest Fock matrix alone.
__global__ void k_diis_solve(const double* G, const long long* count, int nslot, double* coeff, int* status) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  const long long cnt = *count;
  const int n = cnt < nslot ? (int)cnt : nslot;
  const int newest = cnt > 0 ? (int)((cnt - 1) % nslot) : 0;
  for (int i = 0; i < kDiisSlots; ++i) coeff[i] = 0.0;
  coeff[newest] = 1.0;
  *status = 0;
  if (n < 2) return;
  double g = 0.0;
  for (int i = 0; i < n; ++i)
    for (int j = 0; j < n; ++j) {
      const double v = G[i * kDiisSlots + j];
      if (!isfinite(v)) return;
      g = fmax(g, fabs(v));
    }
  const double scale = (isfinite(g) && g > 0.0) ? g : 1.0;
  const int m = n + 1;
  const double ladder[6] = {0.0, 1e-14, 1e-12, 1e-10, 1e-8, 1e-6};
  double A[(kDiisSlots + 1) * (kDiisSlots + 1)], x[kDiisSlots + 1];
  for (int r = 0; r < 6; ++r) {
    for (int i = 0; i < m; ++i) {
      for (int j = 0; j < m; ++j) {
        double v;
        if (i < n && j < n)
          v = G[i * kDiisSlots + j] / scale + (i == j ? ladder[r] : 0.0);
        else if (i == n && j == n)
          v = 0.0;
        else
          v = -1.0;
        A[i * m + j] = v;
      }
      x[i] = (i == n) ? -1.0 : 0.0;
    }
    if (!solve_small(A, x, m)) continue;
    double s = 0.0;
    for (int i = 0; i < n; ++i) s += x[i];
    if (!isfinite(s) || fabs(s) < 1e-15) return;
    for (int i = 0; i < kDiisSlots; ++i) coeff[i] = i < n ? x[i] / s : 0.0;
    *status = 1;
    return;
  }
}

}  // namespace

// sets the message of asuka_jk_last_error (jk_api.hip)
int asuka_fail_msg(int code, const char* msg);

extern "C" {

int asuka_scf_diis_solve(const double* G, const int64_t* count, int nslot, double* coeff, int32_t* status,
                         void* stream) {
  if (!G || !count || !coeff || !status) return asuka_fail_msg(ASUKA_E_INVALID, "G, count, coeff and status must be non-null");
  if (nslot < 1 || nslot > kDiisSlots) return asuka_fail_msg(ASUKA_E_INVALID, "nslot must be in [1, 8]");
  hipLaunchKernelGGL(k_diis_solve, dim3(1), dim3(64), 0, (hipStream_t)stream, G, (const long long*)count, nslot,
                     coeff, (int*)status);
  const hipError_t e = hipGetLastError();
  if (e != hipSuccess) return asuka_fail_msg(ASUKA_E_HIP, (std::string("k_diis_solve: ") + hipGetErrorString(e)).c_str());
  return ASUKA_OK;
}

}  // extern "C"
