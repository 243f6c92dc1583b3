// Generic wave-cooperative Rys ERI kernel with fused J/K (or Fock) digestion.
//
// One wavefront (64 lanes, one workgroup) owns one shell quartet at a time.
// Per primitive quartet:
//   lanes u < nroots      : Rys root/weight u from the interpolation table (LDS)
//   lanes (u, axis) < 3n  : 2-D VRR table G[u][axis][i][k] (LDS), weight and
//                           prefactor folded into the z axis
//   all lanes             : 1-D HRR tables I[u][axis][ia][ib][ic][id] (LDS) via
//                           the binomial transfer from G
//   all lanes             : each lane accumulates its Cartesian components
//                           sum_u Ix*Iy*Iz from LDS
// After the primitive loop the lane's components are digested into J/K with the
// ordered-contraction semantics of the reference
// (cueri_cuda_kernels_dense_scatter.cu:246-352), folding the two symmetric
// updates of each pair into one (the outputs are symmetrised afterwards).
//
// This kernel handles every class up to (ff|ff) with runtime angular momenta;
// the class-specialised register kernels cover the hot low-l classes.
#include "jk_internal.h"
#include "rys_eval.h"

namespace {

constexpr int kWave = 64;
constexpr int kNW = 8;  // components per lane per pass

// Cartesian exponents of the components of l = 0..3 concatenated (offset
// l(l+1)(l+2)/6), component i at bits 2i of an immediate.
constexpr unsigned long long kLX = 0x56B01604ull, kLY = 0x1B18418410ull, kLZ = 0xE491091040ull;
static_assert(((kLX >> 20) & 3) == 3 && ((kLY >> 32) & 3) == 3 && ((kLZ >> 38) & 3) == 3, "exponents");
__device__ __forceinline__ int c_exp(unsigned long long t, int i) { return (int)((t >> (2 * i)) & 3ull); }

// Binomial coefficients C(n, m), n, m <= 3, as 2-bit fields of an immediate
// (bits 2(4n + m)): indexed per lane, a shift and a mask replace a memory load.
constexpr unsigned int kBinom2 = 0x7D190501u;
static_assert(((kBinom2 >> 26) & 3) == 3 && ((kBinom2 >> 18) & 3) == 2 && ((kBinom2 >> 30) & 3) == 1, "C(n, m)");
__device__ __forceinline__ double binom3(int n, int m) { return (double)((kBinom2 >> (2 * (4 * n + m))) & 3u); }

__device__ inline int cart_off(int l) { return l * (l + 1) * (l + 2) / 6; }

__device__ inline double sel3(int ax, double x, double y, double z) { return ax == 0 ? x : (ax == 1 ? y : z); }

__device__ inline double ipow(double x, int k) {
  double r = 1.0;
  for (int i = 0; i < k; ++i) r *= x;
  return r;
}

// At most 128 VGPRs (4 waves per SIMD): the kernel is latency-bound, occupancy
// matters more than the few spilled registers.
__global__ __launch_bounds__(kWave) __attribute__((amdgpu_waves_per_eu(4))) void eri_generic_kernel(DevBasis b, ClassWork w, EriOut o) {
  extern __shared__ double lds[];
  const int lane = threadIdx.x;
  double qmax = 0.0;
  int qmax_pair = -1;

  for (long long k = blockIdx.x; k < w.nwork; k += gridDim.x) {
    int P, Qp;
    if (!decode_work(w, k, P, Qp)) continue;  // segment padding (uniform across the wave)
    if (o.mode <= MODE_FOCK && o.eps_density > 0.0) {
      const double bound = (b.Q[P] * b.Q[Qp]) * fmax(b.Dsp[P], b.Dsp[Qp]);
      if (bound < o.eps_density) continue;  // uniform across the wave
    }
    const int A = b.sp_A[P], B = b.sp_B[P], C = b.sp_A[Qp], Dsh = b.sp_B[Qp];
    const int la = b.shell_l[A], lb = b.shell_l[B], lc = b.shell_l[C], ld = b.shell_l[Dsh];
    const int nA = ncart_c(la), nB = ncart_c(lb), nC = ncart_c(lc), nD = ncart_c(ld);
    const int nCD = nC * nD, ncomp = nA * nB * nCD;
    const int nr = (la + lb + lc + ld) / 2 + 1;
    const int nmax = la + lb, mmax = lc + ld, g1 = (nmax + 1) * (mmax + 1);
    const int sIc = ld + 1, sIb = (lc + 1) * sIc, sIa = (lb + 1) * sIb, n1 = (la + 1) * sIa;

    const double Ax = b.shell_xyz[3 * A], Ay = b.shell_xyz[3 * A + 1], Az = b.shell_xyz[3 * A + 2];
    const double Bx = b.shell_xyz[3 * B], By = b.shell_xyz[3 * B + 1], Bz = b.shell_xyz[3 * B + 2];
    const double Cx = b.shell_xyz[3 * C], Cy = b.shell_xyz[3 * C + 1], Cz = b.shell_xyz[3 * C + 2];
    const double Dx = b.shell_xyz[3 * Dsh], Dy = b.shell_xyz[3 * Dsh + 1], Dz = b.shell_xyz[3 * Dsh + 2];
    const double ABx = Ax - Bx, ABy = Ay - By, ABz = Az - Bz;
    const double CDx = Cx - Dx, CDy = Cy - Dy, CDz = Cz - Dz;

    double* rw = lds;                          // [0, 7): roots, [7, 14): weights * prefactor
    double* G = lds + 2 * ASUKA_NROOTS_MAX;    // nr * 3 * g1
    double* I = G + nr * 3 * g1;               // nr * 3 * n1
    // J/K block accumulators and the density blocks they need (JK/Fock modes)
    const int nAB = nA * nB;
    double* Lj_ab = I + nr * 3 * n1;
    double* Lj_cd = Lj_ab + nAB;
    double* Lk_ac = Lj_cd + nCD;
    double* Lk_ad = Lk_ac + nA * nC;
    double* Lk_bc = Lk_ad + nA * nD;
    double* Lk_bd = Lk_bc + nB * nC;
    double* Dcd = Lk_bd + nB * nD;
    double* Dab = Dcd + nCD;
    double* Dbd = Dab + nAB;
    double* Dbc = Dbd + nB * nD;
    double* Dad = Dbc + nB * nC;
    double* Dac = Dad + nA * nD;
    const int nblk = nAB + nCD + nA * nC + nA * nD + nB * nC + nB * nD;
    const bool digest = o.mode <= MODE_FOCK;
    // two-density builds keep a second set (accumulators, density blocks) of
    // the same layout 2 * nblk further on, filled from D2 and flushed to J2, K2
    const int nden = (digest && o.D2) ? 2 : 1;
    if (digest) {
      long long N, msh;  // the molecule's D/J/K block
      mol_layout(b, A, N, msh);
      const int aoA = b.shell_ao[A], aoB = b.shell_ao[B], aoC = b.shell_ao[C], aoD = b.shell_ao[Dsh];
      for (int di = 0; di < nden; ++di) {
        const double* Dm = (di ? o.D2 : o.D) + msh;
        const int so = di * 2 * nblk;
        for (int e = lane; e < nblk; e += kWave) Lj_ab[so + e] = 0.0;
        for (int e = lane; e < nCD; e += kWave) Dcd[so + e] = Dm[(aoC + e / nD) * N + aoD + e % nD];
        for (int e = lane; e < nAB; e += kWave) Dab[so + e] = Dm[(aoA + e / nB) * N + aoB + e % nB];
        for (int e = lane; e < nB * nD; e += kWave) Dbd[so + e] = Dm[(aoB + e / nD) * N + aoD + e % nD];
        for (int e = lane; e < nB * nC; e += kWave) Dbc[so + e] = Dm[(aoB + e / nC) * N + aoC + e % nC];
        for (int e = lane; e < nA * nD; e += kWave) Dad[so + e] = Dm[(aoA + e / nD) * N + aoD + e % nD];
        for (int e = lane; e < nA * nC; e += kWave) Dac[so + e] = Dm[(aoA + e / nC) * N + aoC + e % nC];
      }
      __syncthreads();
    }

    const int ab0 = b.pp_start[P], nab = b.pp_count[P];
    const int cd0 = b.pp_start[Qp], ncd = b.pp_count[Qp];
    if (o.counters && lane == 0) {
      unsigned long long* cnt = o.counters + kCountStride * w.cp_index;
      atomicAdd(cnt + (((A != B) << 2) | ((C != Dsh) << 1) | (P != Qp)), 1ull);
      atomicAdd(cnt + 8, (unsigned long long)b.pp_count0[P] * (unsigned long long)b.pp_count0[Qp]);
    }
    const int npass = (ncomp + kWave * kNW - 1) / (kWave * kNW);
    const bool single = nab * ncd == 1;
    const int oA = cart_off(la), oB = cart_off(lb), oC = cart_off(lc), oD = cart_off(ld);

    for (int pass = 0; pass < npass; ++pass) {
      int ix[kNW], iy[kNW], iz[kNW];
      double acc[kNW];
#pragma unroll
      for (int j = 0; j < kNW; ++j) {
        acc[j] = 0.0;
        int comp = pass * kWave * kNW + j * kWave + lane;
        if (comp >= ncomp) comp = 0;
        const int iab = comp / nCD, icd = comp - iab * nCD;
        const int ia = iab / nB, ib = iab - ia * nB, ic = icd / nD, id = icd - ic * nD;
        ix[j] = c_exp(kLX, oA + ia) * sIa + c_exp(kLX, oB + ib) * sIb + c_exp(kLX, oC + ic) * sIc + c_exp(kLX, oD + id);
        iy[j] = c_exp(kLY, oA + ia) * sIa + c_exp(kLY, oB + ib) * sIb + c_exp(kLY, oC + ic) * sIc + c_exp(kLY, oD + id);
        iz[j] = c_exp(kLZ, oA + ia) * sIa + c_exp(kLZ, oB + ib) * sIb + c_exp(kLZ, oC + ic) * sIc + c_exp(kLZ, oD + id);
      }

      for (int ipq = 0; ipq < nab * ncd; ++ipq) {
        // with a single primitive quartet (uncontracted shells: every d/f
        // class of these bases) the roots, G and I tables of the first pass
        // stay valid in LDS for the remaining passes
        if (!(single && pass > 0)) {
          const int iab_p = ipq / ncd, icd_p = ipq - iab_p * ncd;
          const PrimPair X = b.pp[ab0 + iab_p];
          const PrimPair Y = b.pp[cd0 + icd_p];
          const double p = X.eta, q = Y.eta;
          const double inv_pq = 1.0 / (p + q);
          const double dPQx = X.Px - Y.Px, dPQy = X.Py - Y.Py, dPQz = X.Pz - Y.Pz;
          const double T = p * q * inv_pq * (dPQx * dPQx + dPQy * dPQy + dPQz * dPQz);
          if (lane < nr) {
            double xr, wr;
            rys_root_dev(nr, lane, T, xr, wr);
            // 2 pi^(5/2) / (p q sqrt(p+q)) cK_ab cK_cd
            const double base = 34.98683665524972497 * X.inv_eta * Y.inv_eta * sqrt(inv_pq) * X.cK * Y.cK;
            rw[lane] = xr;
            rw[ASUKA_NROOTS_MAX + lane] = wr * base;
          }
          __syncthreads();
          if (lane < 3 * nr) {
            const int u = lane / 3, ax = lane - 3 * u;
            const double xr = rw[u];
            const double B0 = xr * 0.5 * inv_pq;
            const double B1 = (1.0 - xr) * 0.5 * X.inv_eta + B0;
            const double B1p = (1.0 - xr) * 0.5 * Y.inv_eta + B0;
            const double Pa = sel3(ax, X.Px, X.Py, X.Pz), Qa = sel3(ax, Y.Px, Y.Py, Y.Pz);
            const double Cc = (Pa - sel3(ax, Ax, Ay, Az)) + (q * inv_pq) * xr * (Qa - Pa);
            const double Cp = (Qa - sel3(ax, Cx, Cy, Cz)) + (p * inv_pq) * xr * (Pa - Qa);
            double* g = G + (u * 3 + ax) * g1;
            const int s = mmax + 1;
            g[0] = (ax == 2) ? rw[ASUKA_NROOTS_MAX + u] : 1.0;
            if (nmax > 0) g[s] = Cc * g[0];
            if (mmax > 0) g[1] = Cp * g[0];
            for (int a = 2; a <= nmax; ++a) g[a * s] = B1 * (double)(a - 1) * g[(a - 2) * s] + Cc * g[(a - 1) * s];
            for (int c = 2; c <= mmax; ++c) g[c] = B1p * (double)(c - 1) * g[c - 2] + Cp * g[c - 1];
            if (nmax > 0 && mmax > 0) {
              for (int a = 1; a <= nmax; ++a) {
                g[a * s + 1] = (double)a * B0 * g[(a - 1) * s] + Cp * g[a * s];
                for (int c = 2; c <= mmax; ++c)
                  g[a * s + c] = B1p * (double)(c - 1) * g[a * s + c - 2] + (double)a * B0 * g[(a - 1) * s + c - 1] +
                                 Cp * g[a * s + c - 1];
              }
            }
          }
          __syncthreads();
          for (int e = lane; e < nr * 3 * n1; e += kWave) {
            const int u = e / (3 * n1);
            const int rem = e - u * 3 * n1;
            const int ax = rem / n1;
            int t = rem - ax * n1;
            const int ia_ = t / sIa;
            t -= ia_ * sIa;
            const int ib_ = t / sIb;
            t -= ib_ * sIb;
            const int ic_ = t / sIc;
            const int id_ = t - ic_ * sIc;
            const double ab = sel3(ax, ABx, ABy, ABz), cd = sel3(ax, CDx, CDy, CDz);
            const double* g = G + (u * 3 + ax) * g1;
            const int s = mmax + 1;
            double val = 0.0;
            for (int n = 0; n <= ib_; ++n) {
              double inner = 0.0;
              for (int m = 0; m <= id_; ++m) inner += binom3(id_, m) * ipow(cd, id_ - m) * g[(ia_ + n) * s + ic_ + m];
              val += binom3(ib_, n) * ipow(ab, ib_ - n) * inner;
            }
            I[e] = val;
          }
          __syncthreads();
        }
#pragma unroll
        for (int j = 0; j < kNW; ++j) {
          double sacc = 0.0;
          for (int u = 0; u < nr; ++u)
            sacc += I[(u * 3) * n1 + ix[j]] * I[(u * 3 + 1) * n1 + iy[j]] * I[(u * 3 + 2) * n1 + iz[j]];
          acc[j] += sacc;
        }
        __syncthreads();
      }

      // ---- digestion of this pass: block sums in LDS ----
      const bool ab_neq = A != B, cd_neq = C != Dsh, bk = P != Qp;
#pragma unroll
      for (int j = 0; j < kNW; ++j) {
        const int comp = pass * kWave * kNW + j * kWave + lane;
        if (comp >= ncomp) continue;
        const double v = acc[j];
        const int iab = comp / nCD, icd = comp - iab * nCD;
        if (o.mode == MODE_TILE) {
          o.tile[w.q_out_off[k] + comp] = v;
          continue;
        }
        if (o.mode == MODE_X3C) {  // ket = (aux shell C, unit s): icd is the aux component
          const int ia = iab / nB, ib = iab - ia * nB;
          const int a = b.shell_ao[A] + ia, bb = b.shell_ao[B] + ib;
          double* Xp = o.X + (long long)(b.shell_ao[C] + icd) * o.x_nao * o.x_nao;
          Xp[(long long)a * o.x_nao + bb] = v;
          if (A != B) Xp[(long long)bb * o.x_nao + a] = v;
          continue;
        }
        if (o.mode == MODE_V2C) {  // bra = (aux shell A, unit s), ket = (aux shell C, unit s)
          o.X[(long long)(b.shell_ao[A] + iab) * o.x_naux + b.shell_ao[C] + icd] = v;
          continue;
        }
        if (o.mode == MODE_SCHWARZ) {
          if (iab == icd && v > qmax) {
            qmax = v;
            qmax_pair = P;
          }
          continue;
        }
        const int ia = iab / nB, ib = iab - ia * nB, ic = icd / nD, id = icd - ic * nD;
        // deterministic builds: the block sums are fixed point with the J and K
        // factors applied here, so the flush adds them unchanged
        const double vj = o.fixed ? v * ((ab_neq ? 2.0 : 1.0) * (cd_neq ? 2.0 : 1.0)) : v;
        const double vk = o.fixed ? v * ((bk ? 2.0 : 1.0) * (o.mode == MODE_FOCK ? -0.5 : 1.0)) : v;
        for (int di = 0; di < nden; ++di) {
          const int so = di * 2 * nblk;
          acc_add(&Lj_ab[so + iab], vj * Dcd[so + icd], o.fixed);
          if (bk) acc_add(&Lj_cd[so + icd], vj * Dab[so + iab], o.fixed);
          acc_add(&Lk_ac[so + ia * nC + ic], vk * Dbd[so + ib * nD + id], o.fixed);
          if (cd_neq) acc_add(&Lk_ad[so + ia * nD + id], vk * Dbc[so + ib * nC + ic], o.fixed);
          if (ab_neq) {
            acc_add(&Lk_bc[so + ib * nC + ic], vk * Dad[so + ia * nD + id], o.fixed);
            if (cd_neq) acc_add(&Lk_bd[so + ib * nD + id], vk * Dac[so + ia * nC + ic], o.fixed);
          }
        }
      }
    }
    if (digest) {
      // flush the block sums: one global atomic per block entry
      __syncthreads();
      long long N, msh;
      mol_layout(b, A, N, msh);
      const int aoA = b.shell_ao[A], aoB = b.shell_ao[B], aoC = b.shell_ao[C], aoD = b.shell_ao[Dsh];
      const bool ab_neq = A != B, cd_neq = C != Dsh, bk = P != Qp;
      const double fj = (ab_neq ? 2.0 : 1.0) * (cd_neq ? 2.0 : 1.0);
      const double fk = (bk ? 2.0 : 1.0) * (o.mode == MODE_FOCK ? -0.5 : 1.0);
      for (int di = 0; di < nden; ++di) {
        double* Jo = shifted(di ? o.J2 : ((o.mode == MODE_FOCK) ? o.F : o.J), msh);
        double* Ko = shifted(di ? o.K2 : ((o.mode == MODE_FOCK) ? o.F : o.K), msh);
        const double* L = Lj_ab + di * 2 * nblk;  // this density's block sums
        const double* Ljab = L;
        const double* Ljcd = Ljab + nAB;
        const double* Lkac = Ljcd + nCD;
        const double* Lkad = Lkac + nA * nC;
        const double* Lkbc = Lkad + nA * nD;
        const double* Lkbd = Lkbc + nB * nC;
        // fixed: the block sums already carry fj / fk (acc_add_sum adds them as they are)
        const double sj = o.fixed ? 1.0 : fj, sk = o.fixed ? 1.0 : fk;
        auto add = [&](double* p, double sum, double sc) {
          if (o.fixed)
            acc_add_sum(p, sum, true);
          else
            atomicAdd(p, sc * sum);
        };
        if (Jo) {
          for (int e = lane; e < nAB; e += kWave) add(&Jo[(aoA + e / nB) * N + aoB + e % nB], Ljab[e], sj);
          if (bk)
            for (int e = lane; e < nCD; e += kWave) add(&Jo[(aoC + e / nD) * N + aoD + e % nD], Ljcd[e], sj);
        }
        if (Ko) {
          for (int e = lane; e < nA * nC; e += kWave) add(&Ko[(aoA + e / nC) * N + aoC + e % nC], Lkac[e], sk);
          if (cd_neq)
            for (int e = lane; e < nA * nD; e += kWave) add(&Ko[(aoA + e / nD) * N + aoD + e % nD], Lkad[e], sk);
          if (ab_neq) {
            for (int e = lane; e < nB * nC; e += kWave) add(&Ko[(aoB + e / nC) * N + aoC + e % nC], Lkbc[e], sk);
            if (cd_neq)
              for (int e = lane; e < nB * nD; e += kWave) add(&Ko[(aoB + e / nD) * N + aoD + e % nD], Lkbd[e], sk);
          }
        }
      }
      __syncthreads();  // LDS is reused by the next quartet
    }
    if (o.mode == MODE_SCHWARZ) {
      // wave max over the lanes' diagonal elements of this quartet
      double m = qmax;
      for (int off = kWave / 2; off > 0; off >>= 1) m = fmax(m, __shfl_xor(m, off));
      if (lane == 0) o.Qout[P] = m > 0.0 ? sqrt(m) : 0.0;
      qmax = 0.0;
      qmax_pair = -1;
    }
  }
  (void)qmax_pair;
}

}  // namespace

size_t eri_generic_lds_bytes(int la, int lb, int lc, int ld, int nden) {
  const int nr = (la + lb + lc + ld) / 2 + 1;
  const int g1 = (la + lb + 1) * (lc + ld + 1);
  const int n1 = (la + 1) * (lb + 1) * (lc + 1) * (ld + 1);
  const int nA = ncart_c(la), nB = ncart_c(lb), nC = ncart_c(lc), nD = ncart_c(ld);
  const int nblk = nA * nB + nC * nD + nA * nC + nA * nD + nB * nC + nB * nD;
  return sizeof(double) * (2 * ASUKA_NROOTS_MAX + nr * 3 * g1 + nr * 3 * n1 + 2 * nden * nblk);
}

hipError_t launch_eri_generic(const DevBasis& b, const ClassWork& w, const EriOut& o, hipStream_t s) {
  if (w.nwork <= 0) return hipSuccess;
  // LDS sized for the class (explicit lists may mix classes: then la..ld hold the maxima)
  const int nden = (o.mode <= MODE_FOCK && o.D2) ? 2 : 1;
  const size_t lds = eri_generic_lds_bytes(w.la, w.lb, w.lc, w.ld, nden);
  if (lds > 64 * 1024) {  // (ff|ff) with two densities
    const hipError_t e =
        hipFuncSetAttribute((const void*)eri_generic_kernel, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    if (e != hipSuccess) return e;
  }
  long long grid = w.nwork < 16384 ? w.nwork : 16384;
  hipLaunchKernelGGL(eri_generic_kernel, dim3((unsigned)grid), dim3(kWave), lds, s, b, w, o);
  return hipGetLastError();
}
