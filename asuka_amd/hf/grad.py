"""Nuclear gradients on the direct path, and J/K as a differentiable op.

* ``direct_jk_grad`` -- backward of the J/K build with respect to the nuclear
  coordinates: d/dR [Tr(W_J J(D)) + Tr(W_K K(D))] at fixed D, W_J, W_K (the
  4c derivative contraction of asuka/mcscf/nac/_dense.py:1513-1839 with the
  adjoint tile of :195-250).
* ``int1e_grad`` -- d/dR [Tr(Dh (T + V)) - Tr(Ws S)] (contract_dhcore_cart
  and contract_dS_ip_cart, asuka/integrals/int1e_cart.py:1545-1738+).
* ``rhf_nuc_grad`` -- the RHF gradient composed as in
  asuka/mcscf/nuc_grad_direct.py:116-183 (there is no stand-alone RHF
  gradient for the direct backend in the reference; its pieces are
  de_h1 + de_2e + de_pulay + de_nuc with D_w = D/2, D_core = D and the
  energy-weighted density W = C_occ diag(2 eps_occ) C_occ^T).
* ``DirectJKFunction`` / ``direct_jk_autograd`` -- ``torch.autograd.Function``
  with forward (J, K) = J/K(D) at the context's geometry R and a hand-written
  backward: dL/dD = J(sym gJ) + K(sym gK), dL/dR = direct_jk_grad(D, sym gJ,
  sym gK).

All device work goes through the C ABI (asuka_jk_grad, asuka_int1e_grad,
asuka_jk_build); there is no CPU fallback.
"""

from __future__ import annotations

import ctypes

import numpy as np

from .. import _lib
from .direct_jk import DirectJKContext, _as_device_matrix, _torch, direct_JK, direct_JK_multi


def _sym(torch, X):
    return 0.5 * (X + X.transpose(0, 1))


def _need_atoms(ctx: DirectJKContext):
    if ctx.natm <= 0:
        raise ValueError("gradients need a context made with atom_coords_bohr and atom_charges")


def direct_jk_grad(ctx: DirectJKContext, D, WJ=None, WK=None):
    """(natm, 3) float64 GPU tensor: d/dR [Tr(WJ J(D)) + Tr(WK K(D))] at fixed
    D, WJ, WK (symmetrised here). For the RHF two-electron energy pass
    WJ = D/2, WK = -D/4."""
    torch = _torch()
    _need_atoms(ctx)
    nao = ctx.nao
    Dd = _as_device_matrix(torch, D, nao, "D", ctx.device)
    wj = None if WJ is None else _sym(torch, _as_device_matrix(torch, WJ, nao, "WJ", ctx.device)).contiguous()
    wk = None if WK is None else _sym(torch, _as_device_matrix(torch, WK, nao, "WK", ctx.device)).contiguous()
    g = torch.empty((ctx.natm, 3), dtype=torch.float64, device=ctx.device)
    _lib.check(_lib.lib().asuka_jk_grad(ctypes.c_void_p(ctx.handle), _lib.ptr(Dd), _lib.ptr(wj), _lib.ptr(wk),
                                        _lib.ptr(g), ctypes.c_void_p(_lib.stream_handle(ctx.device))))
    return g


def int1e_grad(ctx: DirectJKContext, Dh=None, Ws=None):
    """(natm, 3) float64 GPU tensor: d/dR [Tr(Dh (T + V)) - Tr(Ws S)] at fixed
    Dh, Ws (symmetrised here), including the nuclear-attraction operator term."""
    torch = _torch()
    _need_atoms(ctx)
    nao = ctx.nao
    dh = None if Dh is None else _sym(torch, _as_device_matrix(torch, Dh, nao, "Dh", ctx.device)).contiguous()
    ws = None if Ws is None else _sym(torch, _as_device_matrix(torch, Ws, nao, "Ws", ctx.device)).contiguous()
    g = torch.empty((ctx.natm, 3), dtype=torch.float64, device=ctx.device)
    _lib.check(_lib.lib().asuka_int1e_grad(ctypes.c_void_p(ctx.handle), _lib.ptr(dh), _lib.ptr(ws), _lib.ptr(g),
                                           ctypes.c_void_p(_lib.stream_handle(ctx.device))))
    return g


def energy_weighted_density(mo_coeff, mo_energy, mo_occ):
    """W = sum_i n_i eps_i C_i C_i^T (for RHF n_i = 2 on occupied orbitals)."""
    torch = _torch()
    C = mo_coeff if isinstance(mo_coeff, torch.Tensor) else torch.as_tensor(np.asarray(mo_coeff))
    e = mo_energy if isinstance(mo_energy, torch.Tensor) else torch.as_tensor(np.asarray(mo_energy))
    n = mo_occ if isinstance(mo_occ, torch.Tensor) else torch.as_tensor(np.asarray(mo_occ))
    dev = C.device if C.is_cuda else torch.device("cuda", torch.cuda.current_device())
    C = C.to(device=dev, dtype=torch.float64)
    ne = (n.to(device=dev, dtype=torch.float64) * e.to(device=dev, dtype=torch.float64))
    return (C * ne[None, :]) @ C.transpose(0, 1)


def rhf_nuc_grad(run_result, *, parts: dict | None = None):
    """Analytic RHF nuclear gradient (natm, 3) in Eh/Bohr, float64 numpy, of a
    converged ``run_hf_df(..., two_e_backend="direct")`` result, composed as in
    asuka/mcscf/nuc_grad_direct.py:116-183: d(T+V) against D, the 4c derivative
    contraction at D_w = D/2, D_core = D, the Pulay term against
    W = C diag(n eps) C^T, and the nuclear repulsion term.
    ``parts`` (optional) receives the four pieces."""
    torch = _torch()
    ctx = run_result.direct_jk_ctx
    scf = run_result.scf
    _need_atoms(ctx)
    D = scf.dm if scf.dm is not None else None
    C = scf.mo_coeff if isinstance(scf.mo_coeff, torch.Tensor) else torch.as_tensor(np.asarray(scf.mo_coeff))
    occ = scf.mo_occ if isinstance(scf.mo_occ, torch.Tensor) else torch.as_tensor(np.asarray(scf.mo_occ))
    C = C.to(device=ctx.device, dtype=torch.float64)
    occ = occ.to(device=ctx.device, dtype=torch.float64)
    if D is None:
        D = (C * occ[None, :]) @ C.transpose(0, 1)
    D = _as_device_matrix(torch, D, ctx.nao, "D", ctx.device)
    W = energy_weighted_density(C, scf.mo_energy, occ)
    g1 = int1e_grad(ctx, Dh=D, Ws=W)
    g2 = direct_jk_grad(ctx, D, WJ=0.5 * D, WK=-0.25 * D)
    gn = np.asarray(run_result.mol.energy_nuc_grad(), dtype=np.float64)
    g = g1.cpu().numpy() + g2.cpu().numpy() + gn
    if parts is not None:
        parts.update(int1e=g1.cpu().numpy(), two_e=g2.cpu().numpy(), nuc=gn)
    return g


class DirectJKFunction:
    """Namespace for the torch.autograd.Function (built lazily so importing
    this module does not import torch)."""

    _fn = None

    @classmethod
    def get(cls):
        if cls._fn is None:
            torch = _torch()

            class _F(torch.autograd.Function):
                @staticmethod
                def forward(fctx, D, R, jk_ctx, mixed_precision):
                    J, K = direct_JK(jk_ctx, D, mixed_precision=mixed_precision)
                    fctx.jk_ctx = jk_ctx
                    fctx.mixed = mixed_precision
                    fctx.save_for_backward(D.detach())
                    return J, K

                @staticmethod
                def backward(fctx, gJ, gK):
                    (D,) = fctx.saved_tensors
                    jk_ctx = fctx.jk_ctx
                    nao = jk_ctx.nao
                    zero = torch.zeros((nao, nao), dtype=torch.float64, device=D.device)
                    gJ = zero if gJ is None else _sym(torch, gJ.to(torch.float64)).contiguous()
                    gK = zero if gK is None else _sym(torch, gK.to(torch.float64)).contiguous()
                    dD = dR = None
                    if fctx.needs_input_grad[0]:
                        # J and K are linear in D with symmetric kernels: the
                        # adjoint of D -> J(D) is J itself, of D -> K(D) is K.
                        # Both come from one ERI pass of the two-density build.
                        dJ, _, _, dK = direct_JK_multi(jk_ctx, gJ, gK, mixed_precision=fctx.mixed)
                        dD = dJ + dK
                    if fctx.needs_input_grad[1]:
                        dR = direct_jk_grad(jk_ctx, D, WJ=gJ, WK=gK)
                    return dD, dR, None, None

            cls._fn = _F
        return cls._fn


def direct_jk_autograd(jk_ctx: DirectJKContext, D, R, *, mixed_precision: bool | None = None):
    """(J, K) with autograd through D and the nuclear coordinates R (natm, 3),
    which must be the coordinates the context was built at (the context holds
    the basis at that geometry; the backward evaluates the analytic derivative
    there). ``mixed_precision`` applies to the J/K builds of the forward and of
    dL/dD; the derivative contraction for dL/dR is FP64."""
    torch = _torch()
    _need_atoms(jk_ctx)
    R0 = getattr(jk_ctx, "atom_coords_bohr", None)
    if R0 is not None and not np.allclose(np.asarray(R.detach().cpu().numpy()), R0, rtol=0.0, atol=1e-12):
        raise ValueError("R must be the coordinates the context was built at")
    if not isinstance(D, torch.Tensor):
        D = torch.as_tensor(np.asarray(D, dtype=np.float64), device=jk_ctx.device)
    return DirectJKFunction.get().apply(D, R, jk_ctx, mixed_precision)


__all__ = ["DirectJKFunction", "direct_jk_autograd", "direct_jk_grad", "energy_weighted_density", "int1e_grad",
           "rhf_nuc_grad"]
