// C-ABI of the density-fitting integrals: the auxiliary Coulomb metric
// V(P,Q) = (P|Q) and the three-centre integrals X(P, mu, nu) = (mu nu|P).
//
// Replaces metric_2c2e_basis and int3c2e_basis (asuka/cueri/df.py:803-945),
// whose GPU kernels are Rys quadrature with the fourth function a unit s
// (asuka/cueri/cuda/ext/src/cueri_cuda_kernels_rys_generic_part1.cu:1868-2785).
// Here an auxiliary shell P becomes the shell pair (P, u) with u the constant
// function 1 (an s shell of exponent 0 and coefficient 1): its primitive pairs
// are P's primitives (eta = alpha_P, centre R_P, cK = c_P), so (mu nu|P u) =
// (mu nu|P) and (P u|Q u) = (P|Q) run through the same Rys machinery as the
// four-centre integrals (the runtime-class wave-per-quartet kernel) with
// l_d = 0, and the kernel writes the results straight into X or V.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <memory>
#include <string>
#include <vector>

#include "../../include/asuka_jk.h"
#include "jk_internal.h"

int asuka_fail_msg(int code, const char* msg);
hipError_t launch_eri_generic(const DevBasis& b, const ClassWork& w, const EriOut& o, hipStream_t s);

struct asuka_df_ctx {
  int nao = 0, naux = 0;
  int nsh_ao = 0, nsh_aux = 0;
  int nsp_ao = 0;  // AO shell pairs; pair nsp_ao + i is (aux shell i, unit)
  int lmax_ao = 0, lmax_aux = 0;
  DevBasis dev{};
  std::vector<void*> owned;
  ~asuka_df_ctx() {
    for (void* p : owned) (void)hipFree(p);
  }
};

namespace {

template <typename Tp>
int upload(Tp** dst, const std::vector<Tp>& src, std::vector<void*>& owned) {
  *dst = nullptr;
  const size_t bytes = sizeof(Tp) * std::max<size_t>(1, src.size());
  hipError_t e = hipMalloc((void**)dst, bytes);
  if (e != hipSuccess) return asuka_fail_msg(ASUKA_E_NOMEM, (std::string("hipMalloc: ") + hipGetErrorString(e)).c_str());
  owned.push_back(*dst);
  if (!src.empty()) {
    e = hipMemcpy(*dst, src.data(), sizeof(Tp) * src.size(), hipMemcpyHostToDevice);
    if (e != hipSuccess) return asuka_fail_msg(ASUKA_E_HIP, (std::string("hipMemcpy: ") + hipGetErrorString(e)).c_str());
  }
  return ASUKA_OK;
}

int check_basis(const asuka_basis_desc* bd, const char* what) {
  if (!bd || bd->nshell <= 0 || !bd->shell_cxyz || !bd->shell_prim_start || !bd->shell_nprim || !bd->shell_l ||
      !bd->shell_ao_start || !bd->prim_exp || !bd->prim_coef)
    return asuka_fail_msg(ASUKA_E_INVALID, (std::string(what) + " basis arrays must be non-null").c_str());
  for (int i = 0; i < bd->nshell; ++i) {
    const int l = bd->shell_l[i], ps = bd->shell_prim_start[i], np = bd->shell_nprim[i];
    if (l < 0) return asuka_fail_msg(ASUKA_E_INVALID, "shell_l must be >= 0");
    if (l > ASUKA_LMAX)
      return asuka_fail_msg(ASUKA_E_UNSUPPORTED,
                            (std::string(what) + " shells support only l <= " + std::to_string(ASUKA_LMAX)).c_str());
    if (np <= 0 || ps < 0 || ps + np > bd->nprim)
      return asuka_fail_msg(ASUKA_E_INVALID, "shell primitive range out of bounds");
  }
  return ASUKA_OK;
}

}  // namespace

extern "C" {

int asuka_df_ctx_create(const asuka_basis_desc* ao, const asuka_basis_desc* aux, void* stream, asuka_df_ctx** out) {
  (void)stream;
  if (!out) return asuka_fail_msg(ASUKA_E_INVALID, "out must be non-null");
  *out = nullptr;
  int rc;
  if ((rc = check_basis(ao, "AO"))) return rc;
  if ((rc = check_basis(aux, "auxiliary"))) return rc;
  auto ctx = std::make_unique<asuka_df_ctx>();
  const int nsa = ao->nshell, nsx = aux->nshell;
  ctx->nsh_ao = nsa;
  ctx->nsh_aux = nsx;
  // shells: AO shells, auxiliary shells (shell_ao = auxiliary function index), the unit s shell
  std::vector<double> xyz;
  std::vector<int> sl, sao;
  for (int i = 0; i < nsa; ++i) {
    for (int k = 0; k < 3; ++k) xyz.push_back(ao->shell_cxyz[3 * i + k]);
    sl.push_back(ao->shell_l[i]);
    sao.push_back(ao->shell_ao_start[i]);
    ctx->lmax_ao = std::max(ctx->lmax_ao, ao->shell_l[i]);
    ctx->nao = std::max(ctx->nao, ao->shell_ao_start[i] + ncart_c(ao->shell_l[i]));
  }
  for (int i = 0; i < nsx; ++i) {
    for (int k = 0; k < 3; ++k) xyz.push_back(aux->shell_cxyz[3 * i + k]);
    sl.push_back(aux->shell_l[i]);
    sao.push_back(aux->shell_ao_start[i]);
    ctx->lmax_aux = std::max(ctx->lmax_aux, aux->shell_l[i]);
    ctx->naux = std::max(ctx->naux, aux->shell_ao_start[i] + ncart_c(aux->shell_l[i]));
  }
  const int unit = nsa + nsx;
  for (int k = 0; k < 3; ++k) xyz.push_back(0.0);
  sl.push_back(0);
  sao.push_back(0);
  // shell pairs: the AO pairs with l_A >= l_B (asuka/cueri/shell_pairs.py:72-116), then (P, unit)
  std::vector<int> spA, spB, pps, ppc;
  std::vector<PrimPair> pp;
  for (int i = 0; i < nsa; ++i)
    for (int j = 0; j <= i; ++j) {
      int A = i, B = j;
      if (ao->shell_l[i] < ao->shell_l[j]) std::swap(A, B);
      spA.push_back(A);
      spB.push_back(B);
      pps.push_back((int)pp.size());
      const double* a3 = ao->shell_cxyz + 3 * A;
      const double* b3 = ao->shell_cxyz + 3 * B;
      const double dx = a3[0] - b3[0], dy = a3[1] - b3[1], dz = a3[2] - b3[2];
      const double AB2 = dx * dx + dy * dy + dz * dz;
      for (int ia = 0; ia < ao->shell_nprim[A]; ++ia)
        for (int ib = 0; ib < ao->shell_nprim[B]; ++ib) {
          const int pA = ao->shell_prim_start[A] + ia, pB = ao->shell_prim_start[B] + ib;
          const double al = ao->prim_exp[pA], be = ao->prim_exp[pB];
          const double eta = al + be, ie = 1.0 / eta;
          PrimPair r;
          r.eta = eta;
          r.Px = (al * a3[0] + be * b3[0]) * ie;
          r.Py = (al * a3[1] + be * b3[1]) * ie;
          r.Pz = (al * a3[2] + be * b3[2]) * ie;
          r.cK = ao->prim_coef[pA] * ao->prim_coef[pB] * std::exp(-(al * be) * ie * AB2);
          r.a_exp = al;
          r.b_exp = be;
          r.inv_eta = ie;
          if (r.cK != 0.0) pp.push_back(r);
        }
      ppc.push_back((int)pp.size() - pps.back());
    }
  ctx->nsp_ao = (int)spA.size();
  for (int i = 0; i < nsx; ++i) {
    spA.push_back(nsa + i);
    spB.push_back(unit);
    pps.push_back((int)pp.size());
    for (int k = 0; k < aux->shell_nprim[i]; ++k) {
      const int p = aux->shell_prim_start[i] + k;
      PrimPair r;
      r.eta = aux->prim_exp[p];
      r.Px = aux->shell_cxyz[3 * i];
      r.Py = aux->shell_cxyz[3 * i + 1];
      r.Pz = aux->shell_cxyz[3 * i + 2];
      r.cK = aux->prim_coef[p];
      r.a_exp = aux->prim_exp[p];
      r.b_exp = 0.0;
      r.inv_eta = 1.0 / aux->prim_exp[p];
      if (r.cK != 0.0) pp.push_back(r);
    }
    ppc.push_back((int)pp.size() - pps.back());
  }
  std::vector<void*>& own = ctx->owned;
  double* d_xyz;
  int *d_l, *d_ao, *d_spA, *d_spB, *d_pps, *d_ppc, *d_mol;
  PrimPair* d_pp;
  if ((rc = upload(&d_xyz, xyz, own))) return rc;
  if ((rc = upload(&d_l, sl, own))) return rc;
  if ((rc = upload(&d_ao, sao, own))) return rc;
  if ((rc = upload(&d_spA, spA, own))) return rc;
  if ((rc = upload(&d_spB, spB, own))) return rc;
  if ((rc = upload(&d_pps, pps, own))) return rc;
  if ((rc = upload(&d_ppc, ppc, own))) return rc;
  if ((rc = upload(&d_pp, pp, own))) return rc;
  if ((rc = upload(&d_mol, std::vector<int>(sl.size(), 0), own))) return rc;
  DevBasis& dv = ctx->dev;
  dv.shell_xyz = d_xyz;
  dv.shell_l = d_l;
  dv.shell_ao = d_ao;
  dv.sp_A = d_spA;
  dv.sp_B = d_spB;
  dv.pp_start = d_pps;
  dv.pp_count = d_ppc;
  dv.pp_count0 = d_ppc;
  dv.pp = d_pp;
  dv.shell_mol = d_mol;
  dv.nao = ctx->nao;
  dv.nsp = (int)spA.size();
  *out = ctx.release();
  return ASUKA_OK;
}

int asuka_df_ctx_destroy(asuka_df_ctx* ctx) {
  delete ctx;
  return ASUKA_OK;
}

int asuka_df_ctx_info(const asuka_df_ctx* ctx, int64_t out[4]) {
  if (!ctx || !out) return asuka_fail_msg(ASUKA_E_INVALID, "ctx and out must be non-null");
  out[0] = ctx->nao;
  out[1] = ctx->naux;
  out[2] = ctx->nsp_ao;
  out[3] = ctx->nsh_aux;
  return ASUKA_OK;
}

int asuka_df_metric(asuka_df_ctx* ctx, double* V, void* stream) {
  if (!ctx || !V) return asuka_fail_msg(ASUKA_E_INVALID, "ctx and V must be non-null");
  ClassWork w{};
  w.la = w.lc = ctx->lmax_aux;
  w.lb = w.ld = 0;
  w.q_nket = ctx->nsh_aux;
  w.q_bra0 = w.q_ket0 = ctx->nsp_ao;
  w.nwork = (long long)ctx->nsh_aux * ctx->nsh_aux;
  EriOut o{};
  o.mode = MODE_V2C;
  o.X = V;
  o.x_nao = ctx->nao;
  o.x_naux = ctx->naux;
  const hipError_t e = launch_eri_generic(ctx->dev, w, o, (hipStream_t)stream);
  if (e != hipSuccess) return asuka_fail_msg(ASUKA_E_HIP, (std::string("metric: ") + hipGetErrorString(e)).c_str());
  return ASUKA_OK;
}

int asuka_df_int3c2e(asuka_df_ctx* ctx, double* X, void* stream) {
  if (!ctx || !X) return asuka_fail_msg(ASUKA_E_INVALID, "ctx and X must be non-null");
  ClassWork w{};
  w.la = w.lb = ctx->lmax_ao;
  w.lc = ctx->lmax_aux;
  w.ld = 0;
  w.q_nket = ctx->nsh_aux;
  w.q_bra0 = 0;
  w.q_ket0 = ctx->nsp_ao;
  w.nwork = (long long)ctx->nsp_ao * ctx->nsh_aux;
  EriOut o{};
  o.mode = MODE_X3C;
  o.X = X;
  o.x_nao = ctx->nao;
  o.x_naux = ctx->naux;
  const hipError_t e = launch_eri_generic(ctx->dev, w, o, (hipStream_t)stream);
  if (e != hipSuccess) return asuka_fail_msg(ASUKA_E_HIP, (std::string("int3c2e: ") + hipGetErrorString(e)).c_str());
  return ASUKA_OK;
}

}  // extern "C"
