"""Density fitting on the MI355X engine: auxiliary metric, three-centre
integrals, whitening and DF J/K.

Mirrors the reference's DF path (asuka/cueri/df.py, asuka/hf/df_jk.py), the
default two-electron backend of ``run_hf_df`` (asuka/frontend/_scf_dispatch.py:
68-71):

* ``metric_2c2e_basis(aux_basis)`` (df.py:803) -> V(P, Q) = (P|Q);
* ``int3c2e_basis(ao_basis, aux_basis)`` (df.py:859) -> X(mu, nu, P) = (mu nu|P),
  here computed in the BQ layout X[P][mu][nu] (asuka/hf/df_jk.py:445) and
  returned as the (mu, nu, P) view;
* ``cholesky_metric(V)`` (df.py:946) with the reference's diagonal
  regularisation ladder;
* ``whiten_3c2e(X, L)`` (df.py:994): B = X L^-T;
* ``df_J_from_BQ_D``, ``df_K_from_BQ_D``, ``df_K_from_BQ_Cocc``
  (asuka/hf/df_jk.py:445-688).

The integrals run in ``libasuka_jk.so`` (asuka_df_metric, asuka_df_int3c2e):
an auxiliary shell is paired with the constant function 1, so (mu nu|P) and
(P|Q) are four-centre Rys integrals with l_d = 0. The whitening and the J/K
contractions are dense FP64 matrix products on the device (rocBLAS through
torch). There is no CPU fallback.

Auxiliary basis: the reference takes JKFIT sets from basis_set_exchange
(asuka/cueri/df.py:745 recommended_auxbasis), which is not available here;
``even_tempered_auxbasis`` generates an even-tempered fitting set from the
orbital basis instead, and any explicit basis dict can be passed.
"""

from __future__ import annotations

import ctypes
import math
from dataclasses import dataclass, field
from typing import Any

import numpy as np

from . import _lib
from .basis import BasisCartSoA, named_basis, pack_cart_basis, parse_basis_dict


def _torch():
    import torch

    if not torch.cuda.is_available():
        raise RuntimeError("density fitting on this engine needs a GPU (torch.cuda.is_available() is False)")
    return torch


# ---------------------------------------------------------------------------
# Auxiliary basis
# ---------------------------------------------------------------------------


def even_tempered_auxbasis(orbital_basis, elements, *, beta: float = 2.5, lmax: int = 3) -> dict[str, list]:
    """An even-tempered fitting basis for the orbital basis on ``elements``
    (a basis dict in the explicit ``{element: [[l, [exp, 1.0]], ...]}`` form).

    For each element and each auxiliary l up to min(2 l_max, lmax), the
    exponents of the products of orbital primitives whose angular momenta add
    up to at least l (alpha_i + alpha_j) span [e_lo, e_hi]; the set is
    e_lo beta^k, k = 0, 1, ..., up to the first value beyond e_hi, each an
    uncontracted shell. The same construction as PySCF's even-tempered
    auxiliary sets; the reference's JKFIT sets come from basis_set_exchange,
    absent here."""
    spec = named_basis(orbital_basis) if isinstance(orbital_basis, str) else orbital_basis
    shells = parse_basis_dict(spec, list(elements))
    out: dict[str, list] = {}
    for el, blocks in shells.items():
        prims = [(int(l), float(e)) for l, exps, _ in blocks for e in np.asarray(exps).ravel()]
        lmax_orb = max(l for l, _ in prims)
        entry = []
        for la in range(0, min(2 * lmax_orb, int(lmax)) + 1):
            sums = [e1 + e2 for l1, e1 in prims for l2, e2 in prims if l1 + l2 >= la]
            if not sums:
                continue
            lo, hi = min(sums), max(sums)
            nk = max(1, int(math.floor(math.log(hi / lo) / math.log(beta))) + 2)
            for k in range(nk):
                entry.append([la, [lo * beta ** k, 1.0]])
        out[el] = entry
    return out


def build_aux_basis_cart(mol, auxbasis: Any | None = None, orbital_basis: Any | None = None) -> BasisCartSoA:
    """Packed auxiliary basis: an explicit dict, or (None / "etb") the
    even-tempered set of ``even_tempered_auxbasis`` for the orbital basis."""
    elements = sorted({str(s) for s, _ in mol.atoms_bohr})
    if auxbasis is None or (isinstance(auxbasis, str) and auxbasis.strip().lower() == "etb"):
        auxbasis = even_tempered_auxbasis(orbital_basis if orbital_basis is not None else mol.basis, elements)
    if not isinstance(auxbasis, dict):
        raise NotImplementedError(f"auxiliary basis {auxbasis!r}: pass an explicit basis dict or 'etb'")
    return pack_cart_basis(list(mol.atoms_bohr), parse_basis_dict(auxbasis, elements))


# ---------------------------------------------------------------------------
# Integrals
# ---------------------------------------------------------------------------


@dataclass
class DFContext:
    """Device tables of an AO basis and an auxiliary basis (asuka_df_ctx_create)."""

    nao: int
    naux: int
    device: Any
    handle: int
    _keep: tuple = field(default=(), repr=False)

    def __del__(self):
        h = getattr(self, "handle", 0)
        if h:
            try:
                _lib.lib().asuka_df_ctx_destroy(ctypes.c_void_p(h))
            except Exception:
                pass
            self.handle = 0

    def metric(self, out=None):
        """V(P, Q) = (P|Q), (naux, naux) float64 on the device."""
        torch = _torch()
        V = out if out is not None else torch.empty((self.naux, self.naux), dtype=torch.float64, device=self.device)
        _lib.check(_lib.lib().asuka_df_metric(ctypes.c_void_p(self.handle), _lib.ptr(V),
                                              ctypes.c_void_p(_lib.stream_handle(self.device))))
        return V

    def int3c2e_Q(self, out=None):
        """X[P][mu][nu] = (mu nu|P), (naux, nao, nao) float64 on the device."""
        torch = _torch()
        X = out if out is not None else torch.empty((self.naux, self.nao, self.nao), dtype=torch.float64,
                                                    device=self.device)
        _lib.check(_lib.lib().asuka_df_int3c2e(ctypes.c_void_p(self.handle), _lib.ptr(X),
                                               ctypes.c_void_p(_lib.stream_handle(self.device))))
        return X


def make_df_context(ao_basis, aux_basis) -> DFContext:
    from .hf.direct_jk import _basis_desc

    torch = _torch()
    if int(np.asarray(ao_basis.shell_l).size) == 0 or int(np.asarray(aux_basis.shell_l).size) == 0:
        raise ValueError("AO and auxiliary bases must have at least one shell")
    d_ao, k_ao = _basis_desc(ao_basis)
    d_aux, k_aux = _basis_desc(aux_basis)
    h = ctypes.c_void_p()
    _lib.check(_lib.lib().asuka_df_ctx_create(ctypes.byref(d_ao), ctypes.byref(d_aux),
                                              ctypes.c_void_p(_lib.stream_handle()), ctypes.byref(h)))
    info = np.zeros(4, dtype=np.int64)
    _lib.check(_lib.lib().asuka_df_ctx_info(h, info.ctypes.data_as(_lib._i64p)))
    device = torch.device("cuda", torch.cuda.current_device())
    return DFContext(nao=int(info[0]), naux=int(info[1]), device=device, handle=int(h.value),
                     _keep=(d_ao, k_ao, d_aux, k_aux))


def metric_2c2e_basis(aux_basis, *, ao_basis=None, ctx: DFContext | None = None):
    """V(P, Q) = (P|Q) (asuka/cueri/df.py:803-856), (naux, naux) on the device."""
    if ctx is None:
        ctx = make_df_context(ao_basis if ao_basis is not None else aux_basis, aux_basis)
    return ctx.metric()


def int3c2e_basis(ao_basis, aux_basis, *, ctx: DFContext | None = None):
    """X(mu, nu, P) = (mu nu|P) (asuka/cueri/df.py:859-943): the (nao, nao, naux)
    view of the BQ-layout tensor X[P][mu][nu]."""
    if ctx is None:
        ctx = make_df_context(ao_basis, aux_basis)
    return ctx.int3c2e_Q().permute(1, 2, 0)


def cholesky_metric(V):
    """V = L L^T, lower L (asuka/cueri/df.py:946-991): near-singular metrics are
    regularised by adding max(1e-14 max|V_ii|, 1e-12) x 10^k, k = 0..5, to the
    diagonal until the factorisation succeeds with a finite L."""
    torch = _torch()
    V = V.to(torch.float64).clone()
    if V.ndim != 2 or V.shape[0] != V.shape[1]:
        raise ValueError("V must be a square 2D array")
    n = int(V.shape[0])
    if n == 0:
        return torch.empty((0, 0), dtype=torch.float64, device=V.device)
    scale = float(V.diagonal().abs().max())
    shift0 = max(scale * 1e-14, 1e-12)
    added = 0.0
    for k in range(6):
        target = shift0 * 10.0 ** k
        delta = target - added
        if delta != 0.0:
            V.diagonal().add_(delta)
            added = target
        L, info = torch.linalg.cholesky_ex(V)
        if int(info) == 0 and bool(torch.isfinite(L).all()):
            return L
    raise RuntimeError("cholesky_metric failed after diagonal regularization retries")


def whiten_3c2e(X, L):
    """B = X L^-T (asuka/cueri/df.py:994): for X in the BQ layout (naux, nao, nao)
    B[Q] = sum_P (L^-1)_{QP} X[P], returned in the same layout."""
    torch = _torch()
    naux = int(L.shape[0])
    if X.ndim == 3 and int(X.shape[0]) == naux:
        shape = X.shape
        X2 = X.reshape(naux, -1)
    else:
        raise ValueError("X must be (naux, nao, nao) in the BQ layout")
    B2 = torch.linalg.solve_triangular(L, X2, upper=False)
    return B2.reshape(shape)


# ---------------------------------------------------------------------------
# DF J/K (asuka/hf/df_jk.py)
# ---------------------------------------------------------------------------


def df_J_from_BQ_D(BQ, D):
    """J = sum_Q B_Q (B_Q . D) for BQ (naux, nao, nao) (asuka/hf/df_jk.py:445-470)."""
    naux, nao, nao1 = map(int, BQ.shape)
    if nao != nao1 or tuple(D.shape) != (nao, nao):
        raise ValueError("BQ must be (naux, nao, nao) and D (nao, nao)")
    B2 = BQ.reshape(naux, nao * nao)
    d = B2 @ D.reshape(nao * nao)
    return (B2.T @ d).reshape(nao, nao)


def df_K_from_BQ_D(BQ, D):
    """K_mn = sum_Q sum_ls B_Q,ml D_ls B_Q,ns (asuka/hf/df_jk.py:473-525)."""
    naux, nao, _ = map(int, BQ.shape)
    T = BQ @ D  # (naux, nao, nao): (B_Q D)
    T2 = T.permute(1, 0, 2).reshape(nao, naux * nao).contiguous()
    B2 = BQ.permute(1, 0, 2).reshape(nao, naux * nao).contiguous()
    return T2 @ B2.T


def df_K_from_BQ_Cocc(BQ, C_occ, occ=None):
    """K = sum_Q (B_Q C w)(B_Q C w)^T for D = C diag(occ) C^T
    (asuka/hf/df_jk.py:528-685); occ None means 1 for every column."""
    naux, nao, _ = map(int, BQ.shape)
    Cw = C_occ if occ is None else C_occ * occ.sqrt()[None, :]
    nocc = int(Cw.shape[1])
    Y = BQ @ Cw  # (naux, nao, nocc)
    Y2 = Y.permute(1, 0, 2).reshape(nao, naux * nocc).contiguous()
    return Y2 @ Y2.T


def df_JK(BQ, D, *, want_J: bool = True, want_K: bool = True):
    """(J, K) symmetrised from the whitened BQ tensor (asuka/hf/df_scf.py:516-662)."""
    J = df_J_from_BQ_D(BQ, D) if want_J else None
    K = df_K_from_BQ_D(BQ, D) if want_K else None
    if J is not None:
        J = 0.5 * (J + J.T)
    if K is not None:
        K = 0.5 * (K + K.T)
    return J, K


def df_tensors(ao_basis, aux_basis):
    """(ctx, V, L, BQ): the metric, its Cholesky factor and the whitened
    three-centre tensor of an AO/auxiliary basis pair."""
    ctx = make_df_context(ao_basis, aux_basis)
    V = ctx.metric()
    L = cholesky_metric(0.5 * (V + V.T))
    BQ = whiten_3c2e(ctx.int3c2e_Q(), L)
    return ctx, V, L, BQ


__all__ = ["DFContext", "build_aux_basis_cart", "cholesky_metric", "df_JK", "df_J_from_BQ_D", "df_K_from_BQ_Cocc",
           "df_K_from_BQ_D", "df_tensors", "even_tempered_auxbasis", "int3c2e_basis", "make_df_context",
           "metric_2c2e_basis", "whiten_3c2e"]
