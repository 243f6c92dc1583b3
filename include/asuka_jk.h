/*
 * asuka_jk.h -- C ABI of the MI355X (gfx950) integral-direct four-center J/K engine.
 *
 * Plain C, no exceptions across the ABI, no torch or HIP types in the
 * signatures: streams are passed as `void*` (a hipStream_t, NULL = default
 * stream) and all matrices are caller-owned device pointers to row-major
 * float64 (nao, nao) arrays. Every entry point returns 0 on success or a
 * negative ASUKA_E_* code; asuka_jk_last_error() then holds a message
 * (thread-local).
 *
 * Threading: one context per (device, basis). Calls on one context must be
 * serialised by the caller; distinct contexts are independent.
 *
 * Reference interfaces replaced (asuka = zpengmei/ASUKA):
 *   asuka/hf/direct_jk.py:1143  make_direct_jk_context  -> asuka_jk_ctx_create
 *   asuka/hf/direct_jk.py:1530  direct_JK               -> asuka_jk_build
 *   asuka/hf/direct_jk.py:2205  direct_fock_rhf         -> asuka_fock_build
 *   asuka/hf/direct_jk.py:2716  direct_JK_multi         -> asuka_jk_build_multi
 *   asuka/cueri/screening.py:224 schwarz_shellpairs_device -> asuka_jk_schwarz
 *   asuka/cueri/gpu.py:6512     eri_rys_generic_device (tiles) -> asuka_jk_eri_tiles
 *   asuka/integrals/int1e_cart.py:1423 build_int1e_cart -> asuka_int1e
 *   asuka/cueri/_eri_rys_cpu.pyx:2614 rys_roots_weights_cy -> asuka_rys_roots
 * The binding layer replaced is the pybind11 module
 * asuka/cueri/cuda/ext/src/cueri_cuda_ext.cpp:17-65 (caller-allocated outputs,
 * stream-ordered, errors mapped to exceptions).
 */
#ifndef ASUKA_JK_H
#define ASUKA_JK_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define ASUKA_JK_ABI_VERSION 1

/* status codes */
#define ASUKA_OK 0
#define ASUKA_E_INVALID (-1)      /* bad argument (maps to ValueError)           */
#define ASUKA_E_UNSUPPORTED (-2)  /* angular momentum beyond the build (NotImplementedError) */
#define ASUKA_E_HIP (-3)          /* HIP runtime error (RuntimeError)            */
#define ASUKA_E_NOMEM (-4)        /* device allocation failed                     */

/* build flags */
#define ASUKA_JK_WANT_J 1
#define ASUKA_JK_WANT_K 2
#define ASUKA_JK_MIXED 4 /* per-quartet FP32 recursion below opts.tau_mixed */
#define ASUKA_JK_COUNT 8 /* also count the quartets evaluated (asuka_jk_last_counts) */
#define ASUKA_JK_TIME 16 /* record HIP events around the ERI kernels (asuka_jk_last_timing) */
#define ASUKA_JK_DETERMINISTIC 32 /* bitwise reproducible outputs: every J/K/F contribution is rounded to a
                                     multiple of 2^-44 and summed as a 64-bit integer (order-independent),
                                     converted to FP64 once at the end; |element| must stay below 2^19.
                                     Every quartet runs in FP64 (ASUKA_JK_MIXED is ignored). asuka_jk_build,
                                     asuka_jk_build_noalloc, asuka_jk_build_multi and asuka_fock_build; the
                                     batch builds reject it */

typedef struct asuka_jk_ctx asuka_jk_ctx;

/* Packed Cartesian basis (host pointers), the SoA of asuka/cueri/basis_cart.py:11-77:
 * one shell per contraction column, primitive coefficients with normalisation
 * folded in, AO order = shell order with Cartesian components lx-desc, ly-desc. */
typedef struct {
  int32_t nshell;
  const double* shell_cxyz;        /* (nshell, 3) Bohr */
  const int32_t* shell_prim_start; /* (nshell) */
  const int32_t* shell_nprim;      /* (nshell) */
  const int32_t* shell_l;          /* (nshell) */
  const int32_t* shell_ao_start;   /* (nshell) */
  const int32_t* shell_atom;       /* (nshell) owning atom; may be NULL */
  int32_t nprim;
  const double* prim_exp;  /* (nprim) */
  const double* prim_coef; /* (nprim) */
  int32_t natm;
  const double* atom_xyz;    /* (natm, 3) Bohr; may be NULL when natm == 0 */
  const double* atom_charge; /* (natm); may be NULL when natm == 0 */
} asuka_basis_desc;

/* Context options (asuka/hf/direct_jk.py:1143-1150 keyword arguments). */
typedef struct {
  double eps_schwarz; /* keep quartets with Q_ab*Q_cd >= eps; 0 => Q := 1, no screening */
  double eps_density; /* keep quartets with Q_ab*Q_cd*max(Dmax_ab, Dmax_cd) >= eps_density;
                         < 0 => max(eps_schwarz, 1e-10) (asuka/hf/direct_jk.py:1594-1596); 0 disables */
  double tau_mixed;   /* with ASUKA_JK_MIXED: a wavefront of quartets whose Q_ab*Q_cd*max|D|
                         bounds are all below tau_mixed runs the primitive setup, Rys roots,
                         recursion and transfer in FP32 (J/K accumulation stays FP64);
                         <= 0 => 1e-5 */
  int32_t streams;    /* a build's class-pair launches are spread over this many streams: the
                         caller's and streams - 1 of the context's own, forked from and joined
                         back into the caller's stream by events (stream order and graph capture
                         are kept); <= 0 => 4, at most 8 */
  int32_t reserved[7];
} asuka_jk_opts;

/* ABI version and last error message (thread-local). */
int asuka_jk_version(void);
const char* asuka_jk_last_error(void);

/* Build a context: shell pairs, primitive-pair tables, Schwarz bounds (on the
 * device), angular-momentum-class work lists. The context owns its device
 * buffers. Replaces make_direct_jk_context (asuka/hf/direct_jk.py:1143-1470). */
int asuka_jk_ctx_create(const asuka_basis_desc* basis, const asuka_jk_opts* opts, void* stream, asuka_jk_ctx** out);
int asuka_jk_ctx_destroy(asuka_jk_ctx* ctx);

/* Context facts: out[0]=nao, [1]=nshell, [2]=nsp, [3]=Schwarz-screened canonical
 * quartets, [4]=max_l, [5]=number of class pairs with work, [6]=natm,
 * [7]=primitive pairs kept (nonzero coefficient product). */
int asuka_jk_ctx_info(const asuka_jk_ctx* ctx, int64_t out[8]);

/* J and/or K for a symmetric density D. J and K are overwritten and returned
 * symmetrised, ½(X + Xᵀ) (asuka/hf/direct_jk.py:2190-2195). Stream-ordered, no
 * host synchronisation, no device allocation. Replaces direct_JK
 * (asuka/hf/direct_jk.py:1530-2202). J or K may be NULL when not wanted. */
int asuka_jk_build(asuka_jk_ctx* ctx, const double* D, double* J, double* K, int flags, void* stream);

/* asuka_jk_build with its per-build scratch (the class pairs' work-queue heads
 * and the shell-pair density maxima) in a static device array of 4 MiB per GPU
 * instead of the context's buffers: no device allocation, ever. No counterpart
 * in the reference; for callers that cannot allocate device memory (code under
 * HIP graph capture, serving loops, harnesses that forbid hipMalloc). Limits:
 * at most 2048 class pairs and 522240 shell pairs (larger contexts return
 * ASUKA_E_INVALID); ASUKA_JK_TIME is rejected. The array is shared by every
 * call on one GPU and keeps the last call's contents: calls that overlap on
 * one GPU (any context, any stream) race. */
int asuka_jk_build_noalloc(asuka_jk_ctx* ctx, const double* D, double* J, double* K, int flags, void* stream);

/* (Ja, Ka, Jb, Kb) for two symmetric densities (UHF/ROHF alpha and beta)
 * from one evaluation of each shell quartet, digested with Da into (Ja, Ka)
 * and with Db into (Jb, Kb); screening uses max(|Da|, |Db|) per shell-pair
 * block. Outputs as for asuka_jk_build; with ASUKA_JK_WANT_J both Ja and Jb
 * (with ASUKA_JK_WANT_K both Ka and Kb) must be distinct non-NULL buffers.
 * Replaces direct_JK_multi (asuka/hf/direct_jk.py:2716-3095) and its
 * two-density contraction (cueri_cuda_kernels_dense_scatter.cu:551-714). */
int asuka_jk_build_multi(asuka_jk_ctx* ctx, const double* Da, const double* Db, double* Ja, double* Ka, double* Jb,
                         double* Kb, int flags, void* stream);

/* F = h + J - ½K, symmetrised. F must not alias D. Replaces direct_fock_rhf
 * (asuka/hf/direct_jk.py:2205-2713). */
int asuka_fock_build(asuka_jk_ctx* ctx, const double* D, const double* hcore, double* F, int flags, void* stream);

/* Autotuner: times every class pair's kernel launch on density D (median of
 * reps launches after one warm-up) and, for the class-specialised register
 * kernels, every launch shape of {4, 8} wavefronts per workgroup x {8, 16, 32,
 * 64} wave units (64 ket quartets each) per row task x LDS row accumulators
 * on/off, keeps the fastest shape per class pair and spreads the class pairs
 * over the context's streams by their measured times. Synchronous; allocates
 * and frees scratch J/K. Counterpart of the reference's online hot-class
 * autotune (asuka/cueri/gpu.py:4929-5005, ASUKA_CUERI_HOTCLASS_AUTOTUNE). */
int asuka_jk_autotune(asuka_jk_ctx* ctx, const double* D, int flags, int reps, void* stream);

/* Tuning table, one row of 8 int32 per class pair: la, lb, lc, ld, wavefronts
 * per workgroup, wave units per row task, LDS row accumulators (0/1), tuned
 * launch time in microseconds (0: not tuned). get returns the number of class
 * pairs (rows written: min(cap, that)); set applies the rows whose class
 * exists in ctx (a table tuned on one molecule serves others of the same basis). */
int asuka_jk_tuning_get(const asuka_jk_ctx* ctx, int32_t* rows, int cap);
int asuka_jk_tuning_set(asuka_jk_ctx* ctx, const int32_t* rows, int n, void* stream);

/* Shell pairs (l-ordered, asuka/cueri/shell_pairs.py:72-116) and their Schwarz
 * bounds Q_ab = sqrt(max (ab|ab)) (asuka/cueri/screening.py:224-314), host copies. */
int asuka_jk_shell_pairs(const asuka_jk_ctx* ctx, int32_t* sp_A, int32_t* sp_B, double* Q);

/* Quartet statistics per class pair, for the throughput metrics. Fills up to
 * `cap` rows of 8 int64: la, lb, lc, ld, schwarz-screened quartets, primitive
 * quartets, sum over quartets of primitive-quartets x roots, primitive quartets
 * counted before primitive-pair screening. Returns the number of rows (>= 0)
 * or an error code. */
int asuka_jk_class_stats(const asuka_jk_ctx* ctx, int64_t* rows, int cap);

/* Work summary of a context (asuka_jk_stats): the Schwarz-screened quartet
 * count, and from the last build run with ASUKA_JK_COUNT the quartets
 * evaluated (after density screening), their FP32 share (ASUKA_JK_MIXED),
 * their primitive quartets (counted before primitive-pair screening) and
 * the algorithmic FLOPs by the EFM-v1 convention (ERI + digestion; the FP32
 * part of the ERI FLOPs separately); last_ms is the duration of the ERI
 * launches of the last build run with ASUKA_JK_TIME (0 if none finished). */
typedef struct {
  int64_t class_pairs;
  int64_t schwarz_quartets;
  int64_t quartets;
  int64_t quartets_fp32;
  int64_t prim_quartets;
  int64_t prim_quartets_fp32;
  double flops_efm;
  double flops_efm_fp32;
  double last_ms;
  int64_t reserved[7];
} asuka_jk_stats_out;

/* Replaces the stats/profile dicts of direct_JK (asuka/hf/direct_jk.py:1687-1700,
 * :1958-1989) and the cost model of asuka/hf/direct_jk_cost.py:134-294.
 * Synchronises on stream to read the counters. */
int asuka_jk_stats(asuka_jk_ctx* ctx, asuka_jk_stats_out* out, void* stream);

/* Counters of the last build run with ASUKA_JK_COUNT, one row of 16 int64 per
 * class pair in asuka_jk_class_stats order: [0..7] quartets evaluated after
 * density screening, binned by (A!=B)<<2 | (C!=D)<<1 | (bra pair != ket pair)
 * (the symmetry histogram of asuka/hf/direct_jk_cost.py:134-294), [8] their
 * primitive quartets, [9] quartets evaluated in FP32 (ASUKA_JK_MIXED), [10] their
 * primitive quartets, [11] primitive quartets the class-specialised kernels'
 * loops evaluate after the primitive-quartet cut, [12] the lane slots those
 * loops take (64 x the longest lane of each wave unit; [11] / [12] is their
 * SIMD utilisation), [13..15] 0. Synchronises the stream. Returns the number
 * of rows. */
int asuka_jk_last_counts(asuka_jk_ctx* ctx, int64_t* rows, int cap, void* stream);

/* Milliseconds between the HIP events recorded (on the build's stream) before
 * the first and after the last ERI/digestion kernel of the last build run with
 * ASUKA_JK_TIME, and the number of those kernel launches. Synchronises on the
 * end event. */
int asuka_jk_last_timing(asuka_jk_ctx* ctx, double* ms, int64_t* nlaunch);

/* Test hook: contracted Cartesian ERI tiles for an explicit quartet list
 * (device arrays spab, spcd of length nq). Tile q is written at
 * out + out_off[q] in (nAB, nCD) row-major order, with the shell-pair
 * orientation of asuka_jk_shell_pairs (replaces the tile path of
 * asuka/cueri/gpu.py:6512-6600). */
int asuka_jk_eri_tiles(asuka_jk_ctx* ctx, int64_t nq, const int32_t* spab, const int32_t* spcd, const int64_t* out_off,
                       double* out, int flags, void* stream);

/* One-electron integrals S, T, V (nao x nao, device, any may be NULL)
 * (asuka/integrals/int1e_cart.py:1423-1434). */
int asuka_int1e(asuka_jk_ctx* ctx, double* S, double* T, double* V, void* stream);

/* J/K of nmol independent molecules (one context each) in one call on one
 * stream: ctxs[m], D[m], J[m], K[m] as for asuka_jk_build (J or K may be NULL
 * arrays, or hold NULL entries). The builds are issued in turn, each spread
 * over its context's streams, without host synchronisation, so the whole
 * batch can be captured in one HIP graph.
 * The replacement of the per-molecule loop over direct_JK calls (the batched
 * multi-molecule workload); the reference has no batched entry point. */
/* A batch context: nmol independent molecules (one basis each, in the layout
 * of asuka_jk_ctx_create) in one context. Shell pairs and quartets stay
 * within a molecule, every class pair of the whole batch is one launch fed by
 * its device work queue, and each molecule's D/J/K is a packed block of a
 * context-owned buffer. Use it with asuka_jk_batch_build (the single-molecule
 * entry points reject it). Counterpart of the per-molecule loop over
 * direct_JK (the reference has no batched entry point). */
int asuka_jk_batch_create(int nmol, const asuka_basis_desc* const* bases, const asuka_jk_opts* opts, void* stream,
                          asuka_jk_ctx** out);
/* AO count of each molecule of a batch context (min(cap, nmol) written); returns nmol. */
int asuka_jk_batch_info(const asuka_jk_ctx* ctx, int32_t* mol_nao, int cap);
/* J/K of every molecule of a batch context: D[m], J[m], K[m] are the caller's
 * row-major nao_m x nao_m device matrices (J or K may be NULL when not wanted).
 * The densities are copied into the context's packed blocks, one build covers
 * the batch, and the symmetrised blocks are copied out. Stream-ordered, no host
 * synchronisation or allocation (graph-capturable). */
int asuka_jk_batch_build(asuka_jk_ctx* ctx, const double* const* D, double* const* J, double* const* K, int flags,
                         void* stream);

int asuka_jk_build_batched(int nmol, asuka_jk_ctx* const* ctxs, const double* const* D, double* const* J,
                           double* const* K, int flags, void* stream);

/* Backward pass of the J/K build with respect to the nuclear coordinates:
 * grad (natm x 3, device, overwritten) = dL/dR for
 *   L = Tr(WJ J(D)) + Tr(WK K(D))
 * at fixed D, WJ, WK (nao x nao, symmetric, device; WJ or WK may be NULL),
 * over the context's Schwarz-screened quartets. For the RHF two-electron
 * energy pass WJ = D/2, WK = -D/4. Needs atoms (natm > 0) and the shell->atom
 * map. Replaces the 4c derivative contraction of the reference:
 * asuka/mcscf/nac/_dense.py:1513 grad_2e_ham_dense_eri4c_contracted with the
 * adjoint tile of :195 _bar_tile_meanfield_jk, and the kernel
 * asuka/cueri/_eri_rys_cpu.pyx:4316 (CUDA: cueri/gpu.py:6804). */
int asuka_jk_grad(asuka_jk_ctx* ctx, const double* D, const double* WJ, const double* WK, double* grad,
                  void* stream);

/* One-electron part of the nuclear gradient: grad (natm x 3, device,
 * overwritten) = dL1/dR for L1 = Tr(Dh (T + V)) - Tr(Ws S) at fixed symmetric
 * Dh, Ws (either may be NULL), including the nuclear-attraction operator
 * derivative. Replaces contract_dhcore_cart (asuka/integrals/int1e_cart.py:1738)
 * and contract_dS_ip_cart (:1545) as assembled in
 * asuka/mcscf/nuc_grad_direct.py:151-181. */
int asuka_int1e_grad(asuka_jk_ctx* ctx, const double* Dh, const double* Ws, double* grad, void* stream);

/* DIIS extrapolation weights on the device, for an SCF iteration captured in
 * one HIP graph (no host round trip). G: 8 x 8 row-major Gram matrix
 * <e_i, e_j> of the commutator errors kept in a ring of nslot (1..8) slots
 * (slots 0..n-1 in use, n = min(*count, nslot)); count: device int64, vectors
 * pushed so far (the newest in slot (*count - 1) % nslot). Writes coeff (8
 * doubles, one weight per slot, zero for unused slots, summing to 1) and
 * *status (1: extrapolated; 0: the newest slot alone, where the reference
 * returns the newest Fock matrix). Same bordered system, scaling,
 * regularisation ladder and renormalisation as the reference's
 * DIIS.extrapolate (asuka/hf/df_scf.py:722-835); its least-squares retry,
 * unreachable with a finite G, is not taken. Stream-ordered. */
int asuka_scf_diis_solve(const double* G, const int64_t* count, int nslot, double* coeff, int32_t* status,
                         void* stream);

/* ---- Density fitting: auxiliary metric and three-centre integrals ---- */
typedef struct asuka_df_ctx asuka_df_ctx;

/* Shell and primitive-pair tables of an AO basis and an auxiliary basis (both
 * in the packed layout of asuka_basis_desc, l <= 3; atoms unused), on the
 * device. Replaces the plan/cache set-up of int3c2e_basis and
 * metric_2c2e_basis (asuka/cueri/df.py:803-945). */
int asuka_df_ctx_create(const asuka_basis_desc* ao, const asuka_basis_desc* aux, void* stream, asuka_df_ctx** out);
int asuka_df_ctx_destroy(asuka_df_ctx* ctx);
/* out[0]=nao, [1]=naux, [2]=AO shell pairs, [3]=auxiliary shells */
int asuka_df_ctx_info(const asuka_df_ctx* ctx, int64_t out[4]);

/* V (naux x naux, device, row-major, every element written) = (P|Q), the
 * auxiliary Coulomb metric (asuka/cueri/df.py:803 metric_2c2e_basis).
 * Stream-ordered. */
int asuka_df_metric(asuka_df_ctx* ctx, double* V, void* stream);

/* X (naux x nao x nao, device, row-major, every element written) =
 * (mu nu|P) in the BQ layout X[P][mu][nu] of asuka/hf/df_jk.py:445
 * (asuka/cueri/df.py:859 int3c2e_basis returns the same values as X(mu, nu, P)).
 * Stream-ordered. */
int asuka_df_int3c2e(asuka_df_ctx* ctx, double* X, void* stream);

/* Rys roots (x = t^2, ascending) and weights of the n-point rule at T, from the
 * interpolation table the kernels use, evaluated on the host (n <= 7). */
int asuka_rys_roots(int n, double T, double* x, double* w);

/* Same table evaluated on the device for npts arguments: x, w are (npts, n). */
int asuka_rys_roots_device(int n, int64_t npts, const double* T, double* x, double* w, void* stream);

#ifdef __cplusplus
}
#endif

#endif /* ASUKA_JK_H */
