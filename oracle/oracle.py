"""CPU oracle for the direct J/K path -- TEST INFRASTRUCTURE ONLY.

Only ``tests/``, ``__graft_entry__.smoke()`` and the CPU-baseline leg of
``bench.py`` import this module, and only as the checker / the timed CPU
baseline. The product package never imports it.

Python side of the oracle. Each function restates the reference file:line named
in its docstring; the arithmetic kernels live in ``eri_oracle.c`` and
``int1e_oracle.c`` (built into ``liboracle.so`` by the Makefile here).
"""

from __future__ import annotations

import ctypes
import os
import subprocess
from dataclasses import dataclass

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None

_i32p = ctypes.POINTER(ctypes.c_int32)
_f64p = ctypes.POINTER(ctypes.c_double)


class _OracleBasis(ctypes.Structure):
    _fields_ = [
        ("nshell", ctypes.c_int),
        ("shell_cxyz", _f64p),
        ("shell_prim_start", _i32p),
        ("shell_nprim", _i32p),
        ("shell_l", _i32p),
        ("shell_ao_start", _i32p),
        ("prim_exp", _f64p),
        ("prim_coef", _f64p),
        ("nsp", ctypes.c_int),
        ("sp_A", _i32p),
        ("sp_B", _i32p),
        ("sp_pair_start", _i32p),
        ("pair_eta", _f64p),
        ("pair_Px", _f64p),
        ("pair_Py", _f64p),
        ("pair_Pz", _f64p),
        ("pair_cK", _f64p),
    ]


def build() -> str:
    """Compile liboracle.so with the Makefile next to this file."""
    subprocess.run(["make", "-s", "-C", _HERE], check=True)
    return os.path.join(_HERE, "liboracle.so")


def lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(_HERE, "liboracle.so")
        if not os.path.exists(path):
            build()
        L = ctypes.CDLL(path)
        L.oracle_boys_list.argtypes = [ctypes.c_double, ctypes.c_int, _f64p]
        L.oracle_rys.argtypes = [ctypes.c_int, ctypes.c_double, ctypes.c_int, _f64p, _f64p]
        L.oracle_pair_tables.argtypes = [ctypes.POINTER(_OracleBasis), _f64p, _f64p, _f64p, _f64p, _f64p]
        L.oracle_eri_tile.argtypes = [ctypes.POINTER(_OracleBasis), ctypes.c_int, ctypes.c_int, ctypes.c_int, _f64p]
        L.oracle_eri_tile.restype = ctypes.c_int
        L.oracle_schwarz.argtypes = [ctypes.POINTER(_OracleBasis), ctypes.c_int, ctypes.c_int, _f64p]
        L.oracle_direct_jk.argtypes = [ctypes.POINTER(_OracleBasis), ctypes.c_long, _i32p, _i32p, ctypes.c_int,
                                       _f64p, _f64p, _f64p, ctypes.c_int, ctypes.c_int]
        L.oracle_dense_eri.argtypes = [ctypes.POINTER(_OracleBasis), ctypes.c_int, ctypes.c_int, ctypes.c_int, _f64p]
        L.oracle_int1e.argtypes = [ctypes.c_int, _f64p, _i32p, _i32p, _i32p, _i32p, _f64p, _f64p, ctypes.c_int,
                                   _f64p, _f64p, ctypes.c_int, _f64p, _f64p, _f64p]
        L.oracle_int1e.restype = ctypes.c_int
        L.oracle_eri_deriv.argtypes = [ctypes.POINTER(_OracleBasis), ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                       _f64p, _f64p]
        L.oracle_eri_deriv.restype = ctypes.c_int
        _i64p = ctypes.POINTER(ctypes.c_int64)
        L.oracle_screen_counts.argtypes = [ctypes.c_long, _i64p, ctypes.c_long, _f64p, _f64p, ctypes.c_double,
                                           ctypes.c_double, _i32p, _i32p, _i64p, ctypes.c_int, ctypes.c_int, _i64p]
        L.oracle_grad2e.argtypes = [ctypes.POINTER(_OracleBasis), ctypes.c_int, _f64p, _f64p, _f64p, _i32p,
                                    ctypes.c_int, ctypes.c_int, ctypes.c_int, _f64p]
        L.oracle_grad2e_tasks.argtypes = [ctypes.POINTER(_OracleBasis), ctypes.c_int, _f64p, _f64p, _f64p, _i32p,
                                          ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_long, _i32p, _i32p,
                                          _f64p]
        _LIB = L
    return _LIB


def _p(a, t):
    return a.ctypes.data_as(t) if a is not None else None


# ---------------------------------------------------------------------------
# Boys / Rys
# ---------------------------------------------------------------------------


def boys_list(T: float, m_max: int) -> np.ndarray:
    """F_0..F_mmax(T), _eri_rys_cpu.pyx:187-217."""
    out = np.zeros(m_max + 1)
    lib().oracle_boys_list(float(T), int(m_max), _p(out, _f64p))
    return out


def rys_roots_weights(n: int, T: float, hp: bool = True) -> tuple[np.ndarray, np.ndarray]:
    """Rys roots x = t^2 and weights, _eri_rys_cpu.pyx:219-684 (binary128 for n >= 4 when hp)."""
    r = np.zeros(n)
    w = np.zeros(n)
    lib().oracle_rys(int(n), float(T), int(bool(hp)), _p(r, _f64p), _p(w, _f64p))
    return r, w


# ---------------------------------------------------------------------------
# Shell pairs, pair tables, Schwarz, tasks
# ---------------------------------------------------------------------------


@dataclass
class OracleContext:
    basis: object
    sp_A: np.ndarray
    sp_B: np.ndarray
    sp_npair: np.ndarray
    sp_pair_start: np.ndarray
    pair_eta: np.ndarray
    pair_Px: np.ndarray
    pair_Py: np.ndarray
    pair_Pz: np.ndarray
    pair_cK: np.ndarray
    struct: _OracleBasis = None
    _keep: tuple = ()

    @property
    def nsp(self) -> int:
        return int(self.sp_A.shape[0])

    @property
    def nao(self) -> int:
        return int(self.basis.nao)


def shell_pairs_l_order(basis):
    """All unordered shell pairs (i >= j), oriented so l_A >= l_B
    (cueri/shell_pairs.py:72-116)."""
    L = np.asarray(basis.shell_l)
    npr = np.asarray(basis.shell_nprim)
    A, B, npair = [], [], []
    for i in range(basis.nshell):
        for j in range(i + 1):
            a, b = (j, i) if L[i] < L[j] else (i, j)
            A.append(a)
            B.append(b)
            npair.append(int(npr[a]) * int(npr[b]))
    start = np.zeros(len(A) + 1, dtype=np.int32)
    start[1:] = np.cumsum(npair)
    return (np.asarray(A, np.int32), np.asarray(B, np.int32), np.asarray(npair, np.int32), start)


def make_context(basis) -> OracleContext:
    sp_A, sp_B, sp_npair, sp_start = shell_pairs_l_order(basis)
    tot = int(sp_start[-1])
    ctx = OracleContext(basis, sp_A, sp_B, sp_npair, sp_start, *(np.zeros(tot) for _ in range(5)))
    b = basis
    arrs = (np.ascontiguousarray(b.shell_cxyz), np.ascontiguousarray(b.shell_prim_start),
            np.ascontiguousarray(b.shell_nprim), np.ascontiguousarray(b.shell_l),
            np.ascontiguousarray(b.shell_ao_start), np.ascontiguousarray(b.prim_exp),
            np.ascontiguousarray(b.prim_coef))
    st = _OracleBasis(
        int(b.nshell), _p(arrs[0], _f64p), _p(arrs[1], _i32p), _p(arrs[2], _i32p), _p(arrs[3], _i32p),
        _p(arrs[4], _i32p), _p(arrs[5], _f64p), _p(arrs[6], _f64p), ctx.nsp, _p(sp_A, _i32p), _p(sp_B, _i32p),
        _p(sp_start, _i32p), _p(ctx.pair_eta, _f64p), _p(ctx.pair_Px, _f64p), _p(ctx.pair_Py, _f64p),
        _p(ctx.pair_Pz, _f64p), _p(ctx.pair_cK, _f64p))
    ctx.struct = st
    ctx._keep = arrs
    lib().oracle_pair_tables(ctypes.byref(st), _p(ctx.pair_eta, _f64p), _p(ctx.pair_Px, _f64p),
                             _p(ctx.pair_Py, _f64p), _p(ctx.pair_Pz, _f64p), _p(ctx.pair_cK, _f64p))
    return ctx


def eri_tile(ctx: OracleContext, spAB: int, spCD: int, hp: bool = True) -> np.ndarray:
    """Contracted Cartesian tile (nAB, nCD), _eri_rys_cpu.pyx:1096-1291."""
    L = ctx.basis.shell_l
    la, lb = int(L[ctx.sp_A[spAB]]), int(L[ctx.sp_B[spAB]])
    lc, ld = int(L[ctx.sp_A[spCD]]), int(L[ctx.sp_B[spCD]])
    nc = lambda l: (l + 1) * (l + 2) // 2
    out = np.zeros((nc(la) * nc(lb), nc(lc) * nc(ld)))
    rc = lib().oracle_eri_tile(ctypes.byref(ctx.struct), int(spAB), int(spCD), int(bool(hp)), _p(out, _f64p))
    if rc != 0:
        raise ValueError("oracle tile: l too large")
    return out


def schwarz(ctx: OracleContext, hp: bool = True, threads: int = 0) -> np.ndarray:
    """Q_AB = sqrt(max diag (AB|AB)), _eri_rys_cpu.pyx:2486-2611."""
    Q = np.zeros(ctx.nsp)
    lib().oracle_schwarz(ctypes.byref(ctx.struct), int(bool(hp)), int(threads), _p(Q, _f64p))
    return Q


def tasks_screened_sorted_q(Q: np.ndarray, eps: float) -> tuple[np.ndarray, np.ndarray]:
    """Canonical screened quartets (spCD <= spAB, Q_ab Q_cd >= eps)
    (cueri/tasks.py:49-115 ``build_tasks_screened_sorted_q``)."""
    Q = np.asarray(Q, dtype=np.float64).ravel()
    perm = np.argsort(-Q, kind="stable")
    Qs = Q[perm]
    n_valid = int(np.searchsorted(-Qs, 0.0, side="left"))
    if n_valid == 0:
        return np.empty(0, np.int32), np.empty(0, np.int32)
    Qv = Qs[:n_valid]
    if eps > 0:
        thr = float(eps) / np.maximum(Qv, 1e-300)
        jmax = np.searchsorted(-Qv, -thr, side="right").astype(np.int64)
        jmax = np.minimum(jmax, np.arange(n_valid, dtype=np.int64) + 1)
    else:
        jmax = np.arange(1, n_valid + 1, dtype=np.int64)
    i_arr = np.repeat(np.arange(n_valid, dtype=np.int64), jmax)
    offs = np.zeros(n_valid + 1, dtype=np.int64)
    np.cumsum(jmax, out=offs[1:])
    j_arr = np.arange(int(offs[-1]), dtype=np.int64) - offs[i_arr]
    p32 = perm[:n_valid].astype(np.int32)
    ab, cd = p32[i_arr], p32[j_arr]
    sw = cd > ab
    ab2 = np.where(sw, cd, ab).astype(np.int32)
    cd2 = np.where(sw, ab, cd).astype(np.int32)
    return ab2, cd2


def shell_pair_dmax(ctx: OracleContext, D: np.ndarray) -> np.ndarray:
    """max |D_mu nu| over each shell-pair block, both orientations
    (hf/direct_jk.py:1473-1527 ``_compute_D_sp_max``)."""
    b = ctx.basis
    nc = lambda l: (l + 1) * (l + 2) // 2
    out = np.zeros(ctx.nsp)
    for sp in range(ctx.nsp):
        A, B = int(ctx.sp_A[sp]), int(ctx.sp_B[sp])
        a0, b0 = int(b.shell_ao_start[A]), int(b.shell_ao_start[B])
        na, nb = nc(int(b.shell_l[A])), nc(int(b.shell_l[B]))
        blk = np.abs(D[a0:a0 + na, b0:b0 + nb]).max()
        if A != B:
            blk = max(blk, np.abs(D[b0:b0 + nb, a0:a0 + na]).max())
        out[sp] = blk
    return out


def density_screen(ctx, Q, D, ab, cd, eps_density):
    """CSAM mask Q_ab Q_cd max(Dsp_ab, Dsp_cd) >= eps_density (hf/direct_jk.py:1893-1909)."""
    dsp = shell_pair_dmax(ctx, D)
    val = Q[ab] * Q[cd] * np.maximum(dsp[ab], dsp[cd])
    keep = val >= eps_density
    return ab[keep], cd[keep]


def pair_class(la: int, lb: int) -> int:
    """Pair class of an l-ordered shell pair (la >= lb): la(la+1)/2 + lb."""
    return la * (la + 1) // 2 + lb


def nonzero_prim_pairs(ctx: OracleContext) -> np.ndarray:
    """Primitive pairs per shell pair whose coefficient product cK is not an
    exact zero (the K of the work model, hf/direct_jk_cost.py)."""
    nz = ctx.pair_cK != 0.0
    out = np.zeros(ctx.nsp, dtype=np.int64)
    for sp in range(ctx.nsp):
        s0, n = int(ctx.sp_pair_start[sp]), int(ctx.sp_npair[sp])
        out[sp] = int(nz[s0:s0 + n].sum())
    return out


def screen_counts(ctx: OracleContext, Q: np.ndarray, Dsp: np.ndarray | None, eps_schwarz: float,
                  eps_density: float, nprim: np.ndarray | None = None, threads: int = 0) -> dict:
    """Kept canonical quartets per (bra class, ket class) pair and symmetry bin,
    without materialising the list: the Q-descending enumeration of
    build_tasks_screened_sorted_q (cueri/tasks.py:49-115, hf/direct_jk.py:
    1228-1246) and the CSAM mask of hf/direct_jk.py:1893-1909 (Dsp None: Schwarz
    only). Returns {(la, lb, lc, ld): int64[9]}: bins
    (A!=B)<<2 | (C!=D)<<1 | (bra != ket) with the bra the pair of the higher
    class (the later rank inside one class), and [8] the primitive quartets."""
    b = ctx.basis
    L = np.asarray(b.shell_l)
    la = L[ctx.sp_A].astype(np.int64)
    lb = L[ctx.sp_B].astype(np.int64)
    cls = np.ascontiguousarray(la * (la + 1) // 2 + lb, dtype=np.int32)
    neq = np.ascontiguousarray(ctx.sp_A != ctx.sp_B, dtype=np.int32)
    Q = np.ascontiguousarray(Q, dtype=np.float64)
    perm = np.ascontiguousarray(np.argsort(-Q, kind="stable"), dtype=np.int64)
    n_valid = int(np.searchsorted(-Q[perm], 0.0, side="left"))
    if nprim is None:
        nprim = nonzero_prim_pairs(ctx)
    nprim = np.ascontiguousarray(nprim, dtype=np.int64)
    lmax = int(L.max())
    ncls = pair_class(lmax, lmax) + 1
    out = np.zeros((ncls, ncls, 9), dtype=np.int64)
    dsp = None if Dsp is None else np.ascontiguousarray(Dsp, dtype=np.float64)
    _i64p = ctypes.POINTER(ctypes.c_int64)
    lib().oracle_screen_counts(ctypes.c_long(ctx.nsp), _p(perm, _i64p), ctypes.c_long(n_valid), _p(Q, _f64p),
                               _p(dsp, _f64p), float(eps_schwarz), float(eps_density), _p(cls, _i32p),
                               _p(neq, _i32p), _p(nprim, _i64p), ncls, int(threads), _p(out, _i64p))
    names = {}
    for l1 in range(lmax + 1):
        for l2 in range(l1 + 1):
            names[pair_class(l1, l2)] = (l1, l2)
    res = {}
    for X in range(ncls):
        for Y in range(X + 1):
            if out[X, Y].any():
                res[names[X] + names[Y]] = out[X, Y].copy()
    return res


def direct_jk_tasks(ctx: OracleContext, D: np.ndarray, ab: np.ndarray, cd: np.ndarray, hp: bool = True,
                    threads: int = 0, want_J: bool = True, want_K: bool = True):
    """Direct J/K over an explicit quartet list with the digestion semantics of
    cueri_cuda_kernels_dense_scatter.cu:246-352; symmetrised as
    hf/direct_jk.py:2190-2195."""
    nao = ctx.nao
    D = np.ascontiguousarray(D, dtype=np.float64)
    J = np.zeros((nao, nao)) if want_J else None
    K = np.zeros((nao, nao)) if want_K else None
    ab = np.ascontiguousarray(ab, dtype=np.int32)
    cd = np.ascontiguousarray(cd, dtype=np.int32)
    lib().oracle_direct_jk(ctypes.byref(ctx.struct), int(ab.size), _p(ab, _i32p), _p(cd, _i32p), nao,
                           _p(D, _f64p), _p(J, _f64p), _p(K, _f64p), int(bool(hp)), int(threads))
    if J is not None:
        J = 0.5 * (J + J.T)
    if K is not None:
        K = 0.5 * (K + K.T)
    return J, K


def direct_jk(ctx: OracleContext, D: np.ndarray, eps_schwarz: float = 1e-12, eps_density: float | None = None,
              Q: np.ndarray | None = None, hp: bool = True, threads: int = 0):
    """Reference-semantics direct J/K: Schwarz-screened task list
    (hf/direct_jk.py:1217-1246) then density screening at
    eps_density = max(eps_schwarz, 1e-10) (hf/direct_jk.py:1594-1604)."""
    if Q is None:
        Q = schwarz(ctx, hp=hp, threads=threads) if eps_schwarz > 0 else np.ones(ctx.nsp)
    ab, cd = tasks_screened_sorted_q(Q, eps_schwarz)
    if eps_density is None:
        eps_density = max(eps_schwarz, 1e-10)
    if eps_density > 0:
        ab, cd = density_screen(ctx, Q, D, ab, cd, eps_density)
    return direct_jk_tasks(ctx, D, ab, cd, hp=hp, threads=threads)


# ---------------------------------------------------------------------------
# Density fitting
# ---------------------------------------------------------------------------


def _unit_extended(ao, aux):
    """One basis holding the AO shells, the auxiliary shells and, last, the
    constant function 1 as an s shell of exponent 0 and coefficient 1, so that
    (mu nu|P 1) = (mu nu|P) and (P 1|Q 1) = (P|Q) are four-centre tiles."""
    from types import SimpleNamespace

    nsa, nsx = int(ao.shell_l.size), int(aux.shell_l.size)
    npa = int(ao.prim_exp.size)
    ext = SimpleNamespace(
        shell_cxyz=np.ascontiguousarray(np.vstack([ao.shell_cxyz, aux.shell_cxyz, np.zeros((1, 3))])),
        shell_prim_start=np.ascontiguousarray(np.concatenate([ao.shell_prim_start, aux.shell_prim_start + npa,
                                                              [npa + aux.prim_exp.size]]).astype(np.int32)),
        shell_nprim=np.ascontiguousarray(np.concatenate([ao.shell_nprim, aux.shell_nprim, [1]]).astype(np.int32)),
        shell_l=np.ascontiguousarray(np.concatenate([ao.shell_l, aux.shell_l, [0]]).astype(np.int32)),
        shell_ao_start=np.ascontiguousarray(np.concatenate([ao.shell_ao_start, aux.shell_ao_start,
                                                            [0]]).astype(np.int32)),
        prim_exp=np.ascontiguousarray(np.concatenate([ao.prim_exp, aux.prim_exp, [0.0]])),
        prim_coef=np.ascontiguousarray(np.concatenate([ao.prim_coef, aux.prim_coef, [1.0]])),
        nshell=nsa + nsx + 1, nao=0)
    return ext, nsa, nsx


def df_integrals(ao, aux, hp: bool = True):
    """(X, V): X[P][mu][nu] = (mu nu|P) and V[P][Q] = (P|Q) of the reference's
    int3c2e_basis / metric_2c2e_basis (cueri/df.py:803-943), as oracle tiles
    (_eri_rys_cpu.pyx:1096-1291) with the fourth function the constant 1."""
    ext, nsa, nsx = _unit_extended(ao, aux)
    ctx = make_context(ext)
    unit = nsa + nsx
    nc = lambda l: (l + 1) * (l + 2) // 2
    nao = int(ao.shell_ao_start[-1]) + nc(int(ao.shell_l[-1]))
    naux = int(aux.shell_ao_start[-1]) + nc(int(aux.shell_l[-1]))
    pair_of = lambda i, j: (i * (i + 1) // 2 + j) if i >= j else (j * (j + 1) // 2 + i)
    L = ext.shell_l
    X = np.zeros((naux, nao, nao))
    V = np.zeros((naux, naux))
    for p in range(nsx):
        pu = pair_of(unit, nsa + p)
        p0 = int(aux.shell_ao_start[p])
        for i in range(nsa):
            for j in range(i + 1):
                ab = pair_of(i, j)
                A, B = int(ctx.sp_A[ab]), int(ctx.sp_B[ab])
                t = eri_tile(ctx, ab, pu, hp=hp)  # (nA nB, nP)
                a0, b0 = int(ext.shell_ao_start[A]), int(ext.shell_ao_start[B])
                na, nb = nc(int(L[A])), nc(int(L[B]))
                blk = t.reshape(na, nb, -1).transpose(2, 0, 1)
                X[p0:p0 + blk.shape[0], a0:a0 + na, b0:b0 + nb] = blk
                X[p0:p0 + blk.shape[0], b0:b0 + nb, a0:a0 + na] = blk.transpose(0, 2, 1)
        for q in range(nsx):
            t = eri_tile(ctx, pu, pair_of(unit, nsa + q), hp=hp)
            q0 = int(aux.shell_ao_start[q])
            V[p0:p0 + t.shape[0], q0:q0 + t.shape[1]] = t
    return X, V


def df_cholesky(V):
    """Lower L with L L^T = V + s I, s from the reference's ladder
    max(1e-14 max|V_ii|, 1e-12) x 10^k (cueri/df.py:946-991)."""
    V = np.array(V, dtype=np.float64)
    shift0 = max(float(np.abs(np.diag(V)).max()) * 1e-14, 1e-12)
    added = 0.0
    for k in range(6):
        target = shift0 * 10.0 ** k
        V[np.diag_indices_from(V)] += target - added
        added = target
        try:
            L = np.linalg.cholesky(V)
        except np.linalg.LinAlgError:
            continue
        if np.all(np.isfinite(L)):
            return L
    raise RuntimeError("cholesky failed")


def df_JK(X, V, D):
    """DF J and K (hf/df_jk.py:445-525) from B = L^-1 X (cueri/df.py:994):
    J = sum_Q B_Q (B_Q . D), K = sum_Q B_Q D B_Q^T, symmetrised."""
    import scipy.linalg

    naux, nao, _ = X.shape
    L = df_cholesky(0.5 * (V + V.T))
    B = scipy.linalg.solve_triangular(L, X.reshape(naux, -1), lower=True).reshape(naux, nao, nao)
    J = np.einsum("qmn,q->mn", B, np.einsum("qls,ls->q", B, D))
    K = np.einsum("qml,ls,qns->mn", B, D, B, optimize=True)
    return 0.5 * (J + J.T), 0.5 * (K + K.T)


def dense_eri(ctx: OracleContext, hp: bool = True, threads: int = 0) -> np.ndarray:
    """eri_mat[pq, rs] (nao^2 x nao^2), cueri/dense_cpu.py:663-733 semantics."""
    nao = ctx.nao
    eri = np.zeros((nao * nao, nao * nao))
    lib().oracle_dense_eri(ctypes.byref(ctx.struct), nao, int(bool(hp)), int(threads), _p(eri, _f64p))
    return eri


def dense_JK(eri_mat: np.ndarray, D: np.ndarray):
    """hf/dense_jk.py:55-81: J = einsum(mnls,ls), K = einsum(mlns,ls), symmetrised."""
    nao = D.shape[0]
    e4 = eri_mat.reshape(nao, nao, nao, nao)
    J = np.einsum("mnls,ls->mn", e4, D, optimize=True)
    K = np.einsum("mlns,ls->mn", e4, D, optimize=True)
    return 0.5 * (J + J.T), 0.5 * (K + K.T)


# ---------------------------------------------------------------------------
# Nuclear derivatives of the ERIs
# ---------------------------------------------------------------------------


def eri_deriv(ctx: OracleContext, spAB: int, spCD: int, bar: np.ndarray, hp: bool = True) -> np.ndarray:
    """(4, 3) = sum bar * d(ab|cd)/dR_center for centres [A, B, C, D]
    (_eri_rys_cpu.pyx:4316-4656)."""
    bar = np.ascontiguousarray(bar, dtype=np.float64)
    out = np.zeros(12)
    lib().oracle_eri_deriv(ctypes.byref(ctx.struct), int(spAB), int(spCD), int(bool(hp)), _p(bar, _f64p),
                           _p(out, _f64p))
    return out.reshape(4, 3)


def grad2e(ctx: OracleContext, D: np.ndarray, WJ: np.ndarray | None, WK: np.ndarray | None,
           shell_atom: np.ndarray, natm: int, hp: bool = True, threads: int = 0) -> np.ndarray:
    """d/dR [Tr(WJ J(D)) + Tr(WK K(D))] at fixed D, WJ, WK with the reference's
    driver (mcscf/nac/_dense.py:1513-1839: all ordered shell-pair pairs, adjoint
    tile of :195-250 with D_w -> WJ for the Coulomb part and -2 WK for the
    exchange part, D_core -> D)."""
    D = np.ascontiguousarray(D, dtype=np.float64)
    wj = None if WJ is None else np.ascontiguousarray(WJ, dtype=np.float64)
    wk = None if WK is None else np.ascontiguousarray(-2.0 * np.asarray(WK), dtype=np.float64)
    sa = np.ascontiguousarray(shell_atom, dtype=np.int32)
    g = np.zeros((int(natm), 3))
    lib().oracle_grad2e(ctypes.byref(ctx.struct), ctx.nao, _p(wj, _f64p), _p(wk, _f64p), _p(D, _f64p),
                        _p(sa, _i32p), int(natm), int(bool(hp)), int(threads), _p(g, _f64p))
    return g


def grad2e_tasks(ctx: OracleContext, D: np.ndarray, WJ: np.ndarray | None, WK: np.ndarray | None,
                 shell_atom: np.ndarray, natm: int, ab: np.ndarray, cd: np.ndarray, hp: bool = True,
                 threads: int = 0) -> np.ndarray:
    """:func:`grad2e`'s per-quartet contraction over the given (spAB, spCD)
    tasks only (a bounded sample, for timing the CPU baseline)."""
    D = np.ascontiguousarray(D, dtype=np.float64)
    wj = None if WJ is None else np.ascontiguousarray(WJ, dtype=np.float64)
    wk = None if WK is None else np.ascontiguousarray(-2.0 * np.asarray(WK), dtype=np.float64)
    sa = np.ascontiguousarray(shell_atom, dtype=np.int32)
    ab = np.ascontiguousarray(ab, dtype=np.int32)
    cd = np.ascontiguousarray(cd, dtype=np.int32)
    g = np.zeros((int(natm), 3))
    lib().oracle_grad2e_tasks(ctypes.byref(ctx.struct), ctx.nao, _p(wj, _f64p), _p(wk, _f64p), _p(D, _f64p),
                              _p(sa, _i32p), int(natm), int(bool(hp)), int(threads), ctypes.c_long(ab.size),
                              _p(ab, _i32p), _p(cd, _i32p), _p(g, _f64p))
    return g


# ---------------------------------------------------------------------------
# One-electron integrals and the dense RHF
# ---------------------------------------------------------------------------


def int1e(basis, coords_bohr: np.ndarray, charges: np.ndarray):
    """(S, T, V), integrals/int1e_cart.py:1423-1434."""
    b = basis
    nao = int(b.nao)
    S, T, V = (np.zeros((nao, nao)) for _ in range(3))
    xyz = np.ascontiguousarray(coords_bohr, dtype=np.float64)
    Z = np.ascontiguousarray(charges, dtype=np.float64)
    arrs = [np.ascontiguousarray(a) for a in (b.shell_cxyz, b.shell_prim_start, b.shell_nprim, b.shell_l,
                                               b.shell_ao_start, b.prim_exp, b.prim_coef)]
    rc = lib().oracle_int1e(int(b.nshell), _p(arrs[0], _f64p), _p(arrs[1], _i32p), _p(arrs[2], _i32p),
                            _p(arrs[3], _i32p), _p(arrs[4], _i32p), _p(arrs[5], _f64p), _p(arrs[6], _f64p),
                            int(xyz.shape[0]), _p(xyz, _f64p), _p(Z, _f64p), nao, _p(S, _f64p), _p(T, _f64p),
                            _p(V, _f64p))
    if rc != 0:
        raise ValueError("oracle int1e: l too large")
    return S, T, V


class _DIIS:
    """Pulay DIIS, hf/df_scf.py:722-835 (Gram matrix scaled by its abs max,
    regularisation ladder, affine renormalisation)."""

    def __init__(self, max_vec: int = 8):
        self.max_vec = max_vec
        self.F, self.e = [], []

    def push(self, F, e):
        self.F.append(F)
        self.e.append(e)
        if len(self.F) > self.max_vec:
            self.F.pop(0)
            self.e.pop(0)

    def extrapolate(self):
        n = len(self.F)
        if n < 2:
            return self.F[-1]
        E = np.stack([x.ravel() for x in self.e])
        G = E @ E.T
        g = float(np.max(np.abs(G)))
        G = G / (g if g > 0 and np.isfinite(g) else 1.0)
        B = np.zeros((n + 1, n + 1))
        B[:n, :n] = G
        B[:n, n] = B[n, :n] = -1.0
        rhs = np.zeros(n + 1)
        rhs[n] = -1.0
        coeff = None
        for lam in (0.0, 1e-14, 1e-12, 1e-10, 1e-8, 1e-6):
            Br = B.copy()
            Br[:n, :n] += lam * np.eye(n)
            try:
                c = np.linalg.solve(Br, rhs)
            except np.linalg.LinAlgError:
                continue
            if np.all(np.isfinite(c)):
                coeff = c
                break
        if coeff is None:
            return self.F[-1]
        c = coeff[:n]
        s = c.sum()
        if not np.isfinite(s) or abs(s) < 1e-15:
            return self.F[-1]
        c = c / s
        return np.tensordot(c, np.stack(self.F), axes=(0, 0))


def rhf(S, hcore, fock_fn, nelec: int, enuc: float = 0.0, max_cycle: int = 50, conv_tol: float = 1e-10,
        conv_tol_dm: float = 1e-8, diis_space: int = 8, init_fock_cycles: int = 1, dm0=None):
    """RHF loop of hf/direct_scf.py:230-402 with helpers hf/df_scf.py:411-513,
    :665-670; fock_fn(D) -> F = sym(h + J - K/2)."""
    sym = lambda A: 0.5 * (A + A.T)
    s, U = np.linalg.eigh(sym(S))
    X = U @ np.diag(s ** -0.5) @ U.T
    nocc = nelec // 2

    def gen_eigh(F):
        e, Cp = np.linalg.eigh(sym(X.T @ sym(F) @ X))
        return e, X @ Cp

    def density(C):
        Cw = C[:, :nocc] * np.sqrt(2.0)
        return sym(Cw @ Cw.T)

    eps, C = gen_eigh(hcore)
    D = sym(np.asarray(dm0)) if dm0 is not None else density(C)
    if dm0 is None:
        for _ in range(init_fock_cycles):
            eps, C = gen_eigh(fock_fn(D))
            D = density(C)
    diis = _DIIS(diis_space)
    e_last = None
    conv = False
    cycle = 0
    F = None
    for cycle in range(1, max_cycle + 1):
        F = fock_fn(D)
        M = F @ D @ S
        diis.push(F, M - M.T)
        Ft = diis.extrapolate()
        if Ft is None or not np.all(np.isfinite(Ft)):
            diis = _DIIS(diis_space)
        else:
            F = sym(Ft)
        eps, C = gen_eigh(F)
        Dn = density(C)
        e_tot = 0.5 * float(np.sum(Dn * (hcore + F))) + enuc
        dm_err = float(np.linalg.norm(Dn - D))
        de = abs(e_tot - e_last) if e_last is not None else float("inf")
        D, e_last = Dn, e_tot
        if de < conv_tol and dm_err < conv_tol_dm:
            conv = True
            break
    return {"e_tot": e_last, "converged": conv, "niter": cycle, "mo_energy": eps, "mo_coeff": C, "dm": D}


def open_shell_scf(S, hcore, jk_fn, nalpha: int, nbeta: int, enuc: float = 0.0, rohf: bool = False,
                   max_cycle: int = 100, conv_tol: float = 1e-12, conv_tol_dm: float = 1e-10, diis_space: int = 8,
                   diis_start_cycle: int = 2):
    """UHF (hf/direct_scf.py:405-573) or ROHF (:576-) with jk_fn(D) -> (J, K):
    core guess, spin Fock matrices Fs = h + J(Da + Db) - K(Ds), ROHF through
    Roothaan's effective Fock matrix (hf/df_scf.py:691-719), DIIS on
    F D S - (F D S)^T (per spin for UHF, on D_tot for ROHF), and the
    reference's energy expression with this cycle's J and K and the new
    densities (hf/direct_scf.py:543-548)."""
    sym = lambda A: 0.5 * (A + A.T)
    s, U = np.linalg.eigh(sym(S))
    X = U @ np.diag(s ** -0.5) @ U.T
    nao = S.shape[0]

    def gen_eigh(F):
        e, Cp = np.linalg.eigh(sym(X.T @ sym(F) @ X))
        return e, X @ Cp

    def dens(C, n):
        return sym(C[:, :n] @ C[:, :n].T)

    e0, C = gen_eigh(hcore)
    Da, Db = dens(C, nalpha), dens(C, nbeta)
    nd = 1 if rohf else 2
    diis = [_DIIS(diis_space) for _ in range(nd)]
    e_last, conv, cycle = None, False, 0
    for cycle in range(1, max_cycle + 1):
        J, Ka = jk_fn(Da + Db)[0], jk_fn(Da)[1]
        Kb = jk_fn(Db)[1]
        Fa, Fb = sym(hcore + J - Ka), sym(hcore + J - Kb)
        if rohf:
            fc = 0.5 * (Fa + Fb)
            pc, po, pv = Db @ S, (Da - Db) @ S, np.eye(nao) - Da @ S
            G = 0.5 * (pc.T @ fc @ pc + po.T @ fc @ po + pv.T @ fc @ pv) + po.T @ Fb @ pc + po.T @ Fa @ pv \
                + pv.T @ fc @ pc
            Fs, Ds = [G + G.T], [Da + Db]
        else:
            Fs, Ds = [Fa, Fb], [Da, Db]
        if cycle >= diis_start_cycle:
            for k in range(nd):
                M = Fs[k] @ Ds[k] @ S
                diis[k].push(Fs[k], M - M.T)
                Ft = diis[k].extrapolate()
                if Ft is None or not np.all(np.isfinite(Ft)):
                    diis[k] = _DIIS(diis_space)
                else:
                    Fs[k] = sym(Ft)
        ea, Ca = gen_eigh(Fs[0])
        eb, Cb = (ea, Ca) if rohf else gen_eigh(Fs[1])
        Dan, Dbn = dens(Ca, nalpha), dens(Cb, nbeta)
        Dt = Dan + Dbn
        e_tot = float(np.sum(Dt * hcore) + 0.5 * np.sum(Dt * J) - 0.5 * np.sum(Dan * Ka) - 0.5 * np.sum(Dbn * Kb)) \
            + enuc
        dm_err = float(np.linalg.norm(Dan - Da) + np.linalg.norm(Dbn - Db))
        de = abs(e_tot - e_last) if e_last is not None else float("inf")
        Da, Db, e_last = Dan, Dbn, e_tot
        if de < conv_tol and dm_err < conv_tol_dm:
            conv = True
            break
    return {"e_tot": e_last, "converged": conv, "niter": cycle, "dm": (Da, Db)}


__all__ = [
    "boys_list", "build", "dense_JK", "dense_eri", "density_screen", "direct_jk", "direct_jk_tasks", "eri_tile",
    "int1e", "lib", "make_context", "open_shell_scf", "rhf", "rys_roots_weights", "schwarz", "shell_pair_dmax",
    "shell_pairs_l_order", "tasks_screened_sorted_q",
]
