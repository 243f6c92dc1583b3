"""``run_hf_df(mol, two_e_backend="direct")`` on the MI355X engine.

Mirrors the reference's direct branch:
asuka/frontend/scf.py:1219-1277 ``run_hf_df`` ->
asuka/frontend/_scf_dispatch.py:147-184 -> asuka/frontend/scf.py:451-498
``run_rhf_direct`` -> asuka/frontend/_scf_dense_direct_methods.py:321-412
``run_rhf_direct_impl``: closed-shell checks, Cartesian basis and 1e
integrals, direct J/K context, ``rhf_direct``, result with
``two_e_backend="direct"`` and ``direct_jk_ctx``.

Open-shell methods follow the same branch to ``run_uhf_direct`` /
``run_rohf_direct`` (asuka/frontend/_scf_dispatch.py:178-183, asuka/frontend/
scf.py:502-) over the two-density build. Other two-electron backends (DF,
THC, dense) and DFT raise ``NotImplementedError``. One-electron integrals are
computed on the GPU by the same library.
"""

from __future__ import annotations

import os
from dataclasses import dataclass
from typing import Any

import numpy as np

from .basis import build_ao_basis_cart
from .hf.direct_jk import make_direct_jk_context
from .hf.direct_scf import SCFResult, rhf_direct, rohf_direct, uhf_direct


@dataclass(frozen=True)
class Int1eResult:
    S: Any
    T: Any
    V: Any

    @property
    def hcore(self):
        return self.T + self.V


@dataclass(frozen=True)
class RHFDFRunResult:
    """Fields of asuka/frontend/scf.py:170-188 that the direct path fills."""

    mol: Any
    basis_name: str
    auxbasis_name: str
    ao_basis: Any
    aux_basis: Any
    int1e: Int1eResult
    df_B: Any
    scf: SCFResult
    profile: dict | None = None
    two_e_backend: str | None = None
    direct_jk_ctx: Any | None = None


def build_int1e(jk_ctx) -> Int1eResult:
    """S, T, V on the GPU (asuka/integrals/int1e_cart.py:1423-1434)."""
    import ctypes

    import torch

    from . import _lib

    n = jk_ctx.nao
    S, T, V = (torch.empty((n, n), dtype=torch.float64, device=jk_ctx.device) for _ in range(3))
    _lib.check(_lib.lib().asuka_int1e(ctypes.c_void_p(jk_ctx.handle), _lib.ptr(S), _lib.ptr(T), _lib.ptr(V),
                                      ctypes.c_void_p(_lib.stream_handle(jk_ctx.device))))
    return Int1eResult(S=S, T=T, V=V)


def run_rhf_direct(mol, *, basis: Any | None = None, eps_schwarz: float = 1e-12, direct_threads: int = 256,
                   direct_max_tile_bytes: int = 256 << 20, direct_gpu_task_budget_bytes: int | None = None,
                   direct_max_slab_tasks: int | None = None, max_cycle: int = 50, conv_tol: float = 1e-10,
                   conv_tol_dm: float = 1e-8, diis: bool = True, diis_start_cycle: int | None = None,
                   diis_space: int = 8, damping: float = 0.0, level_shift: float = 0.0, dm0=None, mo_coeff0=None,
                   init_fock_cycles: int | None = None, mixed_precision: bool | None = None,
                   profile: dict | None = None, use_graph: bool | None = None,
                   check_every: int = 4) -> RHFDFRunResult:
    if int(getattr(mol, "spin", 0)) != 0:
        raise NotImplementedError("run_rhf_direct currently supports only closed-shell molecules (spin=0)")
    nelec = int(mol.nelectron)
    if nelec <= 0 or nelec % 2 != 0:
        raise ValueError("RHF requires an even, positive electron count")
    if not bool(getattr(mol, "cart", True)):
        raise NotImplementedError("Integral-direct SCF does not yet support spherical AOs (cart=False)")
    ao_basis, basis_name = build_ao_basis_cart(mol, basis)
    jk_ctx = make_direct_jk_context(ao_basis, eps_schwarz=float(eps_schwarz), threads=int(direct_threads),
                                    max_tile_bytes=int(direct_max_tile_bytes),
                                    max_slab_tasks=direct_max_slab_tasks,
                                    gpu_task_budget_bytes=direct_gpu_task_budget_bytes,
                                    atom_coords_bohr=mol.coords_bohr, atom_charges=mol.charges)
    int1e = build_int1e(jk_ctx)
    default_init = int(os.environ.get("ASUKA_HF_INIT_FOCK_CYCLES", "1") or 1)
    nin = default_init if init_fock_cycles is None else max(0, int(init_fock_cycles))
    dstart = int(diis_start_cycle) if diis_start_cycle is not None else (1 if nin > 0 else 2)
    scf = rhf_direct(int1e.S, int1e.hcore, jk_ctx, nelec=nelec, enuc=float(mol.energy_nuc()),
                     max_cycle=int(max_cycle), conv_tol=float(conv_tol), conv_tol_dm=float(conv_tol_dm),
                     diis=bool(diis), diis_start_cycle=dstart, diis_space=int(diis_space), damping=float(damping),
                     level_shift=float(level_shift), dm0=dm0, mo_coeff0=mo_coeff0, init_fock_cycles=nin,
                     mixed_precision=mixed_precision, profile=profile, use_graph=use_graph,
                     check_every=int(check_every))
    return RHFDFRunResult(mol=mol, basis_name=str(basis_name), auxbasis_name="<direct>", ao_basis=ao_basis,
                          aux_basis=None, int1e=int1e, df_B=None, scf=scf, profile=profile, two_e_backend="direct",
                          direct_jk_ctx=jk_ctx)


def run_rhf_df(mol, *, basis: Any | None = None, auxbasis: Any | None = None, max_cycle: int = 50,
               conv_tol: float = 1e-10, conv_tol_dm: float = 1e-8, diis: bool = True,
               diis_start_cycle: int | None = None, diis_space: int = 8, damping: float = 0.0,
               level_shift: float = 0.0, dm0=None, mo_coeff0=None, init_fock_cycles: int | None = None,
               profile: dict | None = None, **_ignored) -> RHFDFRunResult:
    """RHF with density-fitted J/K (the reference's default backend,
    asuka/frontend/_scf_dispatch.py:68-71 -> asuka/hf/df_scf.py:871 rhf_df):
    auxiliary metric and three-centre integrals on the GPU (asuka_amd.df),
    whitened B = L^-1 X, and J/K from B by dense matrix products each cycle.
    ``auxbasis``: an explicit basis dict, or None / "etb" for the even-tempered
    set generated from the orbital basis."""
    from .df import build_aux_basis_cart, df_JK, df_tensors

    if int(getattr(mol, "spin", 0)) != 0:
        raise NotImplementedError("run_rhf_df supports only closed-shell molecules (spin=0)")
    nelec = int(mol.nelectron)
    if nelec <= 0 or nelec % 2 != 0:
        raise ValueError("RHF requires an even, positive electron count")
    ao_basis, basis_name = build_ao_basis_cart(mol, basis)
    aux_basis = build_aux_basis_cart(mol, auxbasis, basis if basis is not None else mol.basis)
    jk_ctx = make_direct_jk_context(ao_basis, eps_schwarz=1e-12, atom_coords_bohr=mol.coords_bohr,
                                    atom_charges=mol.charges)
    int1e = build_int1e(jk_ctx)
    _, _, _, BQ = df_tensors(ao_basis, aux_basis)
    h = int1e.hcore

    def fock(D):
        J, K = df_JK(BQ, D)
        return h + J - 0.5 * K

    default_init = int(os.environ.get("ASUKA_HF_INIT_FOCK_CYCLES", "1") or 1)
    nin = default_init if init_fock_cycles is None else max(0, int(init_fock_cycles))
    dstart = int(diis_start_cycle) if diis_start_cycle is not None else (1 if nin > 0 else 2)
    scf = rhf_direct(int1e.S, h, None, nelec=nelec, enuc=float(mol.energy_nuc()), max_cycle=int(max_cycle),
                     conv_tol=float(conv_tol), conv_tol_dm=float(conv_tol_dm), diis=bool(diis),
                     diis_start_cycle=dstart, diis_space=int(diis_space), damping=float(damping),
                     level_shift=float(level_shift), dm0=dm0, mo_coeff0=mo_coeff0, init_fock_cycles=nin,
                     profile=profile, use_graph=False, fock_fn=fock)
    aux_name = auxbasis if isinstance(auxbasis, str) else ("etb" if auxbasis is None else "<explicit>")
    return RHFDFRunResult(mol=mol, basis_name=str(basis_name), auxbasis_name=str(aux_name), ao_basis=ao_basis,
                          aux_basis=aux_basis, int1e=int1e, df_B=BQ, scf=scf, profile=profile, two_e_backend="df",
                          direct_jk_ctx=None)


def _run_open_shell_direct(kind: str, mol, *, basis: Any | None = None, eps_schwarz: float = 1e-12,
                           direct_threads: int = 256, direct_max_tile_bytes: int = 256 << 20,
                           direct_gpu_task_budget_bytes: int | None = None, direct_max_slab_tasks: int | None = None,
                           max_cycle: int = 50, conv_tol: float = 1e-10, conv_tol_dm: float = 1e-8, diis: bool = True,
                           diis_start_cycle: int = 2, diis_space: int = 8, damping: float = 0.0, dm0=None,
                           mo_coeff0=None, mixed_precision: bool | None = None,
                           profile: dict | None = None) -> RHFDFRunResult:
    nelec = int(mol.nelectron)
    spin = int(getattr(mol, "spin", 0))
    if nelec <= 0 or (nelec + spin) % 2 != 0 or spin < 0 or spin > nelec:
        raise ValueError("nelectron and spin (2S) must give nalpha = (nelec + spin)/2 >= nbeta >= 0")
    if not bool(getattr(mol, "cart", True)):
        raise NotImplementedError("Integral-direct SCF does not yet support spherical AOs (cart=False)")
    nalpha, nbeta = (nelec + spin) // 2, (nelec - spin) // 2
    ao_basis, basis_name = build_ao_basis_cart(mol, basis)
    jk_ctx = make_direct_jk_context(ao_basis, eps_schwarz=float(eps_schwarz), threads=int(direct_threads),
                                    max_tile_bytes=int(direct_max_tile_bytes),
                                    max_slab_tasks=direct_max_slab_tasks,
                                    gpu_task_budget_bytes=direct_gpu_task_budget_bytes,
                                    atom_coords_bohr=mol.coords_bohr, atom_charges=mol.charges)
    int1e = build_int1e(jk_ctx)
    drv = uhf_direct if kind == "uhf" else rohf_direct
    scf = drv(int1e.S, int1e.hcore, jk_ctx, nalpha=nalpha, nbeta=nbeta, enuc=float(mol.energy_nuc()),
              max_cycle=int(max_cycle), conv_tol=float(conv_tol), conv_tol_dm=float(conv_tol_dm), diis=bool(diis),
              diis_start_cycle=int(diis_start_cycle), diis_space=int(diis_space), damping=float(damping), dm0=dm0,
              mo_coeff0=mo_coeff0, mixed_precision=mixed_precision, profile=profile)
    return RHFDFRunResult(mol=mol, basis_name=str(basis_name), auxbasis_name="<direct>", ao_basis=ao_basis,
                          aux_basis=None, int1e=int1e, df_B=None, scf=scf, profile=profile, two_e_backend="direct",
                          direct_jk_ctx=jk_ctx)


def run_uhf_direct(mol, **kwargs) -> RHFDFRunResult:
    """UHF with integral-direct four-centre J/K (asuka/frontend/scf.py:502-560)."""
    return _run_open_shell_direct("uhf", mol, **kwargs)


def run_rohf_direct(mol, **kwargs) -> RHFDFRunResult:
    """ROHF with integral-direct four-centre J/K (asuka/frontend/scf.py:549-)."""
    return _run_open_shell_direct("rohf", mol, **kwargs)


# the open-shell results carry the same fields (asuka/frontend/scf.py UHFDFRunResult, ROHFDFRunResult)
UHFDFRunResult = RHFDFRunResult
ROHFDFRunResult = RHFDFRunResult


def run_hf_df(mol, *, method: str = "rhf", backend: str = "cuda", df: bool = True,
              two_e_backend: str | None = None, guess: Any | None = None, dm0=None, mo_coeff0=None, **kwargs):
    """Unified HF entry point; this engine implements the integral-direct RHF,
    UHF and ROHF branch (asuka/frontend/_scf_dispatch.py:147-184)."""
    method_s = str(method).strip().lower()
    if method_s not in ("rhf", "uhf", "rohf"):
        raise NotImplementedError("only method='rhf', 'uhf' or 'rohf' is implemented by the direct engine")
    backend_s = str(backend).strip().lower()
    if backend_s not in ("cuda", "hip", "gpu"):
        raise ValueError("two_e_backend='direct' requires backend='cuda' (the GPU)")
    teb = str("df" if two_e_backend is None and df else (two_e_backend or "dense")).strip().lower()
    if teb not in ("direct", "df"):
        raise NotImplementedError(f"two_e_backend={teb!r} is not part of this engine ('direct' or 'df')")
    if guess is not None:
        if dm0 is None and hasattr(guess, "dm"):
            dm0 = guess.dm
        elif dm0 is None and isinstance(guess, np.ndarray):
            dm0 = guess
    if teb == "df":
        if method_s != "rhf":
            raise NotImplementedError("the density-fitted backend of this engine implements method='rhf'")
        for k in ("df_layout", "df_threads"):
            kwargs.pop(k, None)
        return run_rhf_df(mol, dm0=dm0, mo_coeff0=mo_coeff0, **kwargs)
    for k in ("auxbasis", "df_layout", "df_threads"):
        kwargs.pop(k, None)
    if method_s == "uhf":
        return run_uhf_direct(mol, dm0=dm0, mo_coeff0=mo_coeff0, **kwargs)
    if method_s == "rohf":
        return run_rohf_direct(mol, dm0=dm0, mo_coeff0=mo_coeff0, **kwargs)
    return run_rhf_direct(mol, dm0=dm0, mo_coeff0=mo_coeff0, **kwargs)


__all__ = ["Int1eResult", "RHFDFRunResult", "ROHFDFRunResult", "UHFDFRunResult", "build_int1e", "run_hf_df",
           "run_rhf_df", "run_rhf_direct", "run_rohf_direct", "run_uhf_direct"]
