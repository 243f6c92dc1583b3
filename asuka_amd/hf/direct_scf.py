"""RHF SCF driven by the direct Fock/J-K engine, on the GPU.

Mirrors ``rhf_direct`` (asuka/hf/direct_scf.py:230-402) and its helpers
(asuka/hf/df_scf.py:411-513 orthogonaliser, generalised eigensolver, RHF
occupations, density; :665-670 commutator error; :722-835 DIIS):

* core-Hamiltonian guess, then ``init_fock_cycles`` predictor cycles;
* per cycle: F from ``direct_fock_rhf`` when max_l <= 2, else
  h + J - K/2 from ``direct_JK`` (asuka/hf/direct_scf.py:331-335);
* DIIS on e = FDS - (FDS)^T from ``diis_start_cycle``, with the reference's
  Gram scaling, regularisation ladder and affine renormalisation, and a reset
  when extrapolation fails;
* eigh through X = S^(-1/2); D = 2 C_occ C_occ^T;
* E = 1/2 sum D o (h + F) + E_nuc; converged when |dE| < conv_tol and
  ||dD||_F < conv_tol_dm.

Everything stays on the device except the few scalars the convergence test
reads. Small eigenproblems (nao <= $ASUKA_HF_CPU_EIGH_MAX_NAO, default 96) are
solved on the host as the reference does (asuka/hf/df_scf.py:27, :424-446).
"""

from __future__ import annotations

import os
import time
from dataclasses import dataclass
from typing import Any

import numpy as np

from .direct_jk import DirectJKContext, _flags, direct_fock_rhf, direct_JK, direct_JK_multi

_CPU_EIGH_MAX_NAO = int(os.environ.get("ASUKA_HF_CPU_EIGH_MAX_NAO", "96") or 96)
# False runs the captured SCF cycle eagerly instead of replaying its graph
# (the same operations; tests compare the two)
_GRAPH_REPLAY = True

# Mixed-precision SCF: a mixed build routes each wave unit of quartets by its
# bound Q_ab Q_cd max|D| against tau, so from one cycle to the next units change
# precision with the density and the Fock matrix moves by up to the FP32 error
# (about 1e-9 at tau = 1e-5 for (H2O)64 cc-pVDZ). That stalls a tight
# convergence test, so once a cycle's ||D_new - D||_F falls below this the
# cycles build in FP64. DIIS keeps its history (restarting it there cost
# tens of cycles at (H2O)64 cc-pVDZ); the switch comes early enough that
# FP64 Fock matrices have replaced the mixed ones in its ring of 8 by
# convergence.
_MIXED_UNTIL_DM = 1e-3


def _mixed_schedule(mixed_precision, mixed_until_dm) -> tuple[bool, float]:
    """(mixed builds at the start, ||D_new - D||_F below which the builds turn
    FP64; 0 keeps them mixed to the end). mixed_until_dm None: the
    environment's ASUKA_SCF_MIXED_UNTIL_DM, else _MIXED_UNTIL_DM."""
    from .. import _lib

    if not (_flags(mixed_precision) & _lib.MIXED):
        return False, 0.0
    if mixed_until_dm is None:
        env = os.environ.get("ASUKA_SCF_MIXED_UNTIL_DM", "").strip()
        mixed_until_dm = float(env) if env else _MIXED_UNTIL_DM
    return True, max(0.0, float(mixed_until_dm))


@dataclass(frozen=True)
class SCFResult:
    """asuka/hf/df_scf.py:838-848."""

    method: str
    converged: bool
    niter: int
    e_tot: float
    e_elec: float
    e_nuc: float
    mo_energy: Any
    mo_coeff: Any
    mo_occ: Any
    dm: Any = None


def _sym(A):
    return 0.5 * (A + A.T)


def _orthogonalizer(torch, S, eps: float = 1e-12):
    s, U = torch.linalg.eigh(_sym(S))
    if bool((s <= eps).any()):
        raise ValueError("S is not positive definite (small/negative eigenvalues)")
    return (U * s.rsqrt()[None, :]) @ U.T


def _gen_eigh(torch, F, X):
    Fp = _sym(X.T @ _sym(F) @ X)
    if Fp.shape[0] <= _CPU_EIGH_MAX_NAO:
        e, Cp = np.linalg.eigh(Fp.cpu().numpy())
        e = torch.as_tensor(e, device=F.device)
        Cp = torch.as_tensor(Cp, device=F.device)
    else:
        e, Cp = torch.linalg.eigh(Fp)
    return e, X @ Cp


def _density(C, nocc: int):
    Cw = C[:, :nocc] * (2.0 ** 0.5)
    return _sym(Cw @ Cw.T)


class _DIIS:
    """asuka/hf/df_scf.py:722-835."""

    def __init__(self, torch, max_vec: int = 8):
        self.torch = torch
        self.max_vec = int(max_vec)
        self.F: list = []
        self.e: list = []

    def push(self, F, e):
        self.F.append(F)
        self.e.append(e)
        if len(self.F) > self.max_vec:
            self.F.pop(0)
            self.e.pop(0)

    def extrapolate(self):
        torch = self.torch
        n = len(self.F)
        if n < 2:
            return self.F[-1]
        E = torch.stack([x.reshape(-1) for x in self.e])
        G = (E @ E.T).cpu().numpy()
        if not np.all(np.isfinite(G)):
            return self.F[-1]
        g = float(np.max(np.abs(G)))
        G = G / (g if np.isfinite(g) and g > 0 else 1.0)
        B = np.zeros((n + 1, n + 1))
        B[:n, :n] = G
        B[:n, n] = B[n, :n] = -1.0
        rhs = np.zeros(n + 1)
        rhs[n] = -1.0
        coeff = None
        ladder = (0.0, 1e-14, 1e-12, 1e-10, 1e-8, 1e-6)
        for lam in ladder:
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
            for lam in ladder[1:]:
                Br = B.copy()
                Br[:n, :n] += lam * np.eye(n)
                c = np.linalg.lstsq(Br, rhs, rcond=None)[0]
                if np.all(np.isfinite(c)):
                    coeff = c
                    break
        if coeff is None:
            return self.F[-1]
        c = coeff[:n]
        s = float(c.sum())
        if not np.isfinite(s) or abs(s) < 1e-15:
            return self.F[-1]
        c = torch.as_tensor(c / s, device=self.F[-1].device, dtype=torch.float64)
        out = torch.tensordot(c, torch.stack(self.F), dims=([0], [0]))
        if not bool(torch.isfinite(out).all()):
            return self.F[-1]
        return out


def rhf_direct(
    S,
    hcore,
    jk_ctx: DirectJKContext,
    *,
    nelec: int,
    enuc: float = 0.0,
    max_cycle: int = 50,
    conv_tol: float = 1e-10,
    conv_tol_dm: float = 1e-8,
    diis: bool = True,
    diis_start_cycle: int = 1,
    diis_space: int = 8,
    damping: float = 0.0,
    level_shift: float = 0.0,
    dm0=None,
    mo_coeff0=None,
    init_fock_cycles: int = 1,
    mixed_precision: bool | None = None,
    profile: dict | None = None,
    use_graph: bool | None = None,
    check_every: int = 4,
    fock_fn=None,
    mixed_until_dm: float | None = None,
) -> SCFResult:
    """RHF with integral-direct J/K (asuka/hf/direct_scf.py:230-402).

    ``use_graph`` (default ``$ASUKA_SCF_GRAPH``, else on when the Fock build
    serves the basis, max l <= 2): every SCF iteration after the predictor
    cycles is one replay of a captured HIP graph (Fock build, DIIS with the
    device solve, density by purification, energy and convergence test), and
    the host reads the convergence record every ``check_every`` iterations
    (:func:`_rhf_graph_iterations`).

    With mixed-precision builds (``mixed_precision``, default
    ``$ASUKA_ERI_MIXED_PRECISION`` or on) the cycles switch to FP64 builds
    once ||D_new - D||_F falls below ``mixed_until_dm`` (default
    ``$ASUKA_SCF_MIXED_UNTIL_DM`` or 1e-3; 0 keeps them mixed): near
    convergence the density-dependent FP32 routing would move the Fock matrix
    by its FP32 error from cycle to cycle."""
    import torch

    dev = jk_ctx.device if jk_ctx is not None else torch.device("cuda", torch.cuda.current_device())
    S = torch.as_tensor(np.asarray(S) if not isinstance(S, torch.Tensor) else S).to(dev, torch.float64)
    h = torch.as_tensor(np.asarray(hcore) if not isinstance(hcore, torch.Tensor) else hcore).to(dev, torch.float64)
    nao = int(S.shape[0])
    if tuple(S.shape) != (nao, nao) or tuple(h.shape) != (nao, nao):
        raise ValueError("S/hcore must be (nao, nao)")
    if jk_ctx is None and fock_fn is None:
        raise ValueError("rhf_direct needs a DirectJKContext or a fock_fn")
    if jk_ctx is not None and nao != jk_ctx.nao:
        raise ValueError(f"nao mismatch: S has {nao}, DirectJKContext has {jk_ctx.nao}")
    nelec = int(nelec)
    if nelec < 0 or nelec % 2:
        raise ValueError("RHF requires an even nelec >= 0")
    nocc = nelec // 2
    occ = torch.zeros(nao, dtype=torch.float64, device=dev)
    occ[:nocc] = 2.0
    X = _orthogonalizer(torch, S)

    use_fock = fock_fn is None and int(jk_ctx.max_l) <= 2
    mixed0, mixed_until = _mixed_schedule(mixed_precision, mixed_until_dm)
    prec = {"mixed": mixed0}  # the builds' precision now (the eager cycles switch it)

    def fock(D):
        if fock_fn is not None:  # another two-electron backend (density fitting)
            return _sym(fock_fn(D))
        if use_fock:
            return direct_fock_rhf(jk_ctx, D, h, mixed_precision=prec["mixed"])
        J, K = direct_JK(jk_ctx, D, mixed_precision=prec["mixed"])
        return _sym(h + J - 0.5 * K)

    def shifted(F):
        if level_shift:
            Fp = X.T @ F @ X + float(level_shift) * torch.eye(nao, dtype=torch.float64, device=dev)
            return X @ Fp @ X.T
        return F

    if use_graph is None:
        env = os.environ.get("ASUKA_SCF_GRAPH", "").strip()
        use_graph = (env != "0") if env else use_fock
    use_graph = bool(use_graph and use_fock and int(diis_start_cycle) <= 1 and int(max_cycle) > 0)

    def guess_density(F):
        """Aufbau density of F: eigh, or with the graph driver purification
        (the same projector, no eigensolver)."""
        if use_graph:
            Fo = X.T @ shifted(F) @ X
            P, _ = _purify(torch, _sym(Fo), nocc, _purify_steps(torch, _sym(Fo), nocc) + 4)
            return None, None, _sym(2.0 * (X @ P @ X.T))
        e, Cm = _gen_eigh(torch, shifted(F), X)
        return e, Cm, _density(Cm, nocc)

    if dm0 is not None:
        D = _sym(torch.as_tensor(np.asarray(dm0) if not isinstance(dm0, torch.Tensor) else dm0).to(dev, torch.float64))
        if tuple(D.shape) != (nao, nao):
            raise ValueError("dm0 must have shape (nao, nao)")
        eps = C = None
    elif mo_coeff0 is not None:
        C = torch.as_tensor(np.asarray(mo_coeff0) if not isinstance(mo_coeff0, torch.Tensor) else mo_coeff0).to(
            dev, torch.float64)
        if tuple(C.shape) != (nao, nao):
            raise ValueError("mo_coeff0 must have shape (nao, nao)")
        D = _density(C, nocc)
        eps = None
    else:
        eps, C, D = guess_density(h)

    lam = float(damping) if damping else 0.0
    F_last = None  # the last Fock matrix built before the cycles (sizes the graph's purification)
    keep = profile is not None and bool(profile.get("keep_densities"))
    if init_fock_cycles > 0 and dm0 is None and mo_coeff0 is None:
        for _ in range(int(init_fock_cycles)):
            Dp = D
            if keep:
                profile.setdefault("scf", {}).setdefault("predictor_densities", []).append(Dp.clone())
            F_last = fock(Dp)
            eps, C, D = guess_density(F_last)
            if lam:
                D = (1.0 - lam) * D + lam * Dp

    if use_graph:
        return _rhf_graph_iterations(torch, S, h, X, D, jk_ctx, F_probe=F_last, nocc=nocc, occ=occ, enuc=float(enuc),
                                     max_cycle=int(max_cycle), conv_tol=float(conv_tol),
                                     conv_tol_dm=float(conv_tol_dm), diis=bool(diis), diis_space=int(diis_space),
                                     damping=lam, level_shift=float(level_shift or 0.0),
                                     mixed_precision=mixed0, mixed_until=mixed_until,
                                     check_every=max(1, int(check_every)), profile=profile)

    diis_obj = _DIIS(torch, diis_space) if diis else None
    e_last = None
    converged = False
    cycle = 0
    njk = 0
    energies, dm_errors = [], []
    fp64_from = None  # first cycle with an FP64 build after the switch
    for cycle in range(1, int(max_cycle) + 1):
        F = shifted(fock(D))
        njk += 1
        if diis_obj is not None and cycle >= int(diis_start_cycle):
            M = F @ D @ S
            diis_obj.push(F, M - M.T)
            try:
                Ft = diis_obj.extrapolate()
            except Exception:
                Ft = None
            if Ft is None or not bool(torch.isfinite(Ft).all()):
                diis_obj = _DIIS(torch, diis_space)
            else:
                F = _sym(Ft)
        eps, C = _gen_eigh(torch, F, X)
        Dn = _density(C, nocc)
        if lam:
            Dn = (1.0 - lam) * Dn + lam * D
        e_tot = 0.5 * float((Dn * (h + F)).sum().item()) + float(enuc)
        dm_err = float(torch.linalg.norm(Dn - D).item())
        de = abs(e_tot - e_last) if e_last is not None else float("inf")
        D, e_last = Dn, e_tot
        energies.append(e_tot)
        dm_errors.append(dm_err)
        if de < conv_tol and dm_err < conv_tol_dm:
            converged = True
            break
        if prec["mixed"] and dm_err < mixed_until:
            prec["mixed"] = False
            fp64_from = cycle + 1
    if profile is not None:
        profile.setdefault("scf", {})["iters"] = int(cycle)
        profile["scf"]["jk_calls"] = njk
        profile["scf"]["fp64_from_cycle"] = fp64_from
        profile["scf"]["graph"] = False
        profile["scf"]["energies"] = energies
        profile["scf"]["dm_errors"] = dm_errors
    return SCFResult(method="RHF", converged=converged, niter=int(cycle),
                     e_tot=float(e_last if e_last is not None else float("nan")),
                     e_elec=float(e_last - float(enuc)) if e_last is not None else float("nan"),
                     e_nuc=float(enuc), mo_energy=eps, mo_coeff=C, mo_occ=occ, dm=D)


_PURIFY_TOL = 1e-11  # ||P^2 - P||_F at which a purification stops updating


def _purify(torch, Fo, nocc: int, iters: int):
    """Projector onto the eigenvectors of the nocc lowest eigenvalues of the
    symmetric Fo, by trace-correcting purification (TC2; A. M. N. Niklasson,
    Phys. Rev. B 66, 155115 (2002)): the Gershgorin interval [e0, e1] of Fo
    maps its spectrum into [0, 1] reversed, P = (e1 - Fo)/(e1 - e0); then each
    step takes P^2 when Tr P exceeds nocc and 2P - P^2 otherwise, which drives
    the nocc largest eigenvalues of P (the nocc lowest of Fo) to 1 and the rest
    to 0 for a nonzero HOMO-LUMO gap. Each map repels one end of [0, 1] (P^2
    doubles the distance of eigenvalues below 1), so steps past convergence
    would undo it: once ||P^2 - P||_F falls below _PURIFY_TOL the projector is
    frozen (a sticky device flag). Matrix products only, and the branch and the
    freeze are device-side selects, so a fixed number of steps can be captured
    in a HIP graph. Returns P and the idempotency error of the returned P."""
    n = int(Fo.shape[0])
    d = torch.diagonal(Fo)
    r = Fo.abs().sum(dim=1) - d.abs()
    e0 = (d - r).min()
    e1 = (d + r).max()
    eye = torch.eye(n, dtype=Fo.dtype, device=Fo.device)
    P = (e1 * eye - Fo) / (e1 - e0)
    done = torch.zeros((), dtype=torch.bool, device=Fo.device)
    err = None
    for _ in range(int(iters)):
        P2 = P @ P
        err = torch.linalg.norm(P2 - P)
        done = done | (err < _PURIFY_TOL)
        step = torch.where(torch.trace(P) > float(nocc), P2, 2.0 * P - P2)
        P = torch.where(done, P, step)
    return _sym(P), err


def _purify_steps(torch, Fo, nocc: int, max_steps: int = 300) -> int:
    """Steps _purify takes on Fo until ||P^2 - P||_F < _PURIFY_TOL (eager, the
    host reads the error every step)."""
    n = int(Fo.shape[0])
    d = torch.diagonal(Fo)
    r = Fo.abs().sum(dim=1) - d.abs()
    e0 = (d - r).min()
    e1 = (d + r).max()
    P = (e1 * torch.eye(n, dtype=Fo.dtype, device=Fo.device) - Fo) / (e1 - e0)
    for k in range(1, max_steps + 1):
        P2 = P @ P
        if float(torch.linalg.norm(P2 - P)) < _PURIFY_TOL:
            return k
        P = P2 if float(torch.trace(P)) > nocc else 2.0 * P - P2
    raise RuntimeError("density purification did not converge (no HOMO-LUMO gap?)")


def _rhf_graph_iterations(torch, S, h, X, D0, jk_ctx, *, F_probe=None, nocc, occ, enuc, max_cycle, conv_tol,
                          conv_tol_dm, diis, diis_space, damping, level_shift, mixed_precision, check_every, profile,
                          mixed_until=0.0):
    """The SCF cycles of rhf_direct as replays of one captured HIP graph.

    One replay is one cycle of asuka/hf/direct_scf.py:337-378 with its state
    in fixed device buffers: the Fock build of D (asuka_fock_build), the DIIS
    push of F and e = FDS - (FDS)^T into a ring of diis_space slots with the
    Gram row <e_i, e> updated in place, the extrapolation weights from
    asuka_scf_diis_solve, the new density 2 X P X^T with P the purified
    projector of X^T F X (_purify, in place of eigh + aufbau: the same
    projector for a gapped spectrum, to the purification's idempotency
    tolerance), the energy 1/2 sum D_new (h + F) + E_nuc, and the convergence
    test |dE| < conv_tol and ||D_new - D||_F < conv_tol_dm, written to a
    per-cycle record. The host replays check_every cycles, reads the record,
    and stops at the first converged cycle, whose density and Fock matrix are
    restored from a ring of the last check_every cycles. Orbitals and orbital
    energies come from one eigh of the final Fock matrix."""
    import ctypes

    from .. import _lib

    dev = S.device
    nao = int(S.shape[0])
    f64 = torch.float64
    nslot = max(1, min(int(diis_space), 8)) if diis else 1
    use_fock_out = torch.empty((nao, nao), dtype=f64, device=dev)

    # purification steps: measured on the last Fock matrix built before the
    # cycles (the predictor's; else one build of D0), with a margin
    F0 = F_probe if F_probe is not None else direct_fock_rhf(jk_ctx, D0, h, mixed_precision=mixed_precision)
    npur = int(1.25 * _purify_steps(torch, _sym(X.T @ F0 @ X), nocc)) + 6
    nbuild = 1 if F_probe is None else 0  # Fock builds outside the cycles made here

    st = {
        "D": D0.clone(), "F": torch.empty((nao, nao), dtype=f64, device=dev),
        "Fs": torch.zeros((8, nao, nao), dtype=f64, device=dev), "Es": torch.zeros((8, nao * nao), dtype=f64, device=dev),
        "G": torch.zeros((8, 8), dtype=f64, device=dev), "count": torch.zeros(1, dtype=torch.int64, device=dev),
        "coeff": torch.zeros(8, dtype=f64, device=dev), "status": torch.zeros(1, dtype=torch.int32, device=dev),
        "E": torch.full((1,), float("nan"), dtype=f64, device=dev), "it": torch.zeros(1, dtype=torch.int64, device=dev),
        "recE": torch.zeros(max_cycle + 1, dtype=f64, device=dev),
        "recC": torch.zeros(max_cycle + 1, dtype=torch.bool, device=dev),
        "recP": torch.zeros(max_cycle + 1, dtype=f64, device=dev),
        "recD": torch.zeros(max_cycle + 1, dtype=f64, device=dev),
        "ringD": torch.zeros((check_every, nao, nao), dtype=f64, device=dev),
        "ringF": torch.zeros((check_every, nao, nao), dtype=f64, device=dev),
    }
    eye = torch.eye(nao, dtype=f64, device=dev)
    L = _lib.lib()

    def cycle(build: bool = True, mixed: bool = bool(mixed_precision)):
        D = st["D"]
        if build:
            F = direct_fock_rhf(jk_ctx, D, h, mixed_precision=mixed, out=use_fock_out)
        else:  # warm-up of the other operations: the Fock build needs none
            F = use_fock_out.copy_(F0)
        if level_shift:
            F = X @ (X.T @ F @ X + level_shift * eye) @ X.T
        if diis:
            M = F @ D @ S
            e = (M - M.T).reshape(-1)
            slot = torch.remainder(st["count"], nslot)
            st["Fs"].index_copy_(0, slot, F[None])
            st["Es"].index_copy_(0, slot, e[None])
            g = st["Es"] @ e
            st["G"].index_copy_(0, slot, g[None])
            st["G"].index_copy_(1, slot, g[:, None])
            st["count"].add_(1)
            _lib.check(L.asuka_scf_diis_solve(_lib.ptr(st["G"]), _lib.ptr(st["count"]), nslot, _lib.ptr(st["coeff"]),
                                              _lib.ptr(st["status"]), ctypes.c_void_p(_lib.stream_handle(dev))))
            F = _sym(torch.tensordot(st["coeff"], st["Fs"], dims=([0], [0])))
        P, perr = _purify(torch, _sym(X.T @ F @ X), nocc, npur)
        Dn = _sym(2.0 * (X @ P @ X.T))
        if damping:
            Dn = (1.0 - damping) * Dn + damping * D
        E = 0.5 * (Dn * (h + F)).sum() + enuc
        dm_err = torch.linalg.norm(Dn - D)
        conv = ((E - st["E"]).abs() < conv_tol) & (dm_err < conv_tol_dm)
        it = st["it"]
        st["recE"].index_copy_(0, it, E.reshape(1))
        st["recC"].index_copy_(0, it, conv.reshape(1))
        st["recP"].index_copy_(0, it, perr.reshape(1))
        st["recD"].index_copy_(0, it, dm_err.reshape(1))
        ring = torch.remainder(it, check_every)
        st["ringD"].index_copy_(0, ring, Dn[None])
        st["ringF"].index_copy_(0, ring, F[None])
        st["E"].copy_(E.reshape(1))
        st["D"].copy_(Dn)
        st["F"].copy_(F)
        it.add_(1)

    # warm the captured operations once outside the graph on a side stream
    # (the allocator's pools and the library's lazy set-up), then restore the state
    saved = {k: v.clone() for k, v in st.items()}
    side = torch.cuda.Stream(device=dev)
    side.wait_stream(torch.cuda.current_stream(dev))
    with torch.cuda.stream(side):
        cycle(build=False)
    torch.cuda.current_stream(dev).wait_stream(side)
    torch.cuda.synchronize(dev)
    for k, v in saved.items():
        st[k].copy_(v)
    del saved
    # mixed builds switching to FP64 (mixed_until > 0): one graph per precision,
    # each with its own memory pool (a shared pool is only safe for graphs
    # replayed in the order they were captured; the FP64 graph replays alone)
    switch = bool(mixed_precision) and mixed_until > 0.0
    graphs = {}
    if _GRAPH_REPLAY:
        for mp in ((True, False) if switch else (bool(mixed_precision),)):
            g = torch.cuda.CUDAGraph()
            with torch.cuda.graph(g):
                cycle(mixed=mp)
            graphs[mp] = g
    mixed_now = bool(mixed_precision)
    fp64_from = None  # first cycle (1-based) with an FP64 build after the switch
    converged = False
    ncyc = 0
    first = None
    # profile["keep_densities"]: keep each cycle's input density (a device copy
    # per cycle, 2 nao^2 doubles of traffic) so that the Fock builds of the
    # cycles can be timed alone on the same densities
    keep = profile is not None and bool(profile.get("keep_densities"))
    dens = [st["D"].clone()] if keep else None
    torch.cuda.synchronize(dev)
    t_loop = time.perf_counter()
    while ncyc < max_cycle and first is None:
        # while the builds are mixed and may switch, the record is read after
        # every cycle, so the switch comes at the cycle the eager driver would
        # take (one scalar read per cycle against a cycle of milliseconds)
        k = 1 if (mixed_now and switch) else min(check_every, max_cycle - ncyc)
        for _ in range(k):
            if graphs:
                graphs[mixed_now].replay()
            else:
                cycle(mixed=mixed_now)
            if keep:
                dens.append(st["D"].clone())
        ncyc += k
        rec = st["recC"][:ncyc].cpu().numpy()
        hits = np.nonzero(rec)[0]
        if hits.size:
            first = int(hits[0])
        elif mixed_now and switch and float(st["recD"][ncyc - 1]) < mixed_until:
            mixed_now = False
            fp64_from = ncyc + 1  # 1-based, as in the eager driver
    t_loop = time.perf_counter() - t_loop  # the replayed cycles (the host's record reads included)
    nrep = ncyc  # cycles replayed (those after the first converged one are discarded)
    if first is not None:
        converged = True
        ncyc = first + 1
        r = first % check_every
        D = st["ringD"][r].clone()
        F = st["ringF"][r].clone()
    else:
        D = st["D"].clone()
        F = st["F"].clone()
    recE = st["recE"][:ncyc].cpu().numpy()
    e_tot = float(recE[-1])
    eps, C = _gen_eigh(torch, F, X)
    if profile is not None:
        profile.setdefault("scf", {})["iters"] = int(ncyc)
        profile["scf"]["jk_calls"] = int(nrep) + nbuild
        profile["scf"]["cycles_replayed"] = int(nrep)
        profile["scf"]["fp64_from_cycle"] = fp64_from
        if keep:
            profile["scf"]["cycle_densities"] = dens[:nrep]
        profile["scf"]["cycle_loop_s"] = t_loop
        profile["scf"]["graph"] = True
        profile["scf"]["purify_steps"] = npur
        profile["scf"]["purify_err_max"] = float(st["recP"][:ncyc].max())
        profile["scf"]["energies"] = recE.tolist()
        profile["scf"]["dm_errors"] = st["recD"][:ncyc].cpu().numpy().tolist()
        profile["scf"]["purify_errors"] = st["recP"][:ncyc].cpu().numpy().tolist()
    return SCFResult(method="RHF", converged=converged, niter=int(ncyc), e_tot=e_tot, e_elec=e_tot - enuc,
                     e_nuc=enuc, mo_energy=eps, mo_coeff=C, mo_occ=occ, dm=D)


def _occ_uhf(nalpha: int, nbeta: int, nao: int):
    """asuka/hf/df_scf.py:473-484."""
    nalpha, nbeta = int(nalpha), int(nbeta)
    if nalpha < 0 or nbeta < 0:
        raise ValueError("nalpha/nbeta must be >= 0")
    if nalpha > nao or nbeta > nao:
        raise ValueError("nalpha/nbeta exceeds number of orbitals")
    return nalpha, nbeta


def _density_spin(C, nocc: int):
    Co = C[:, :nocc]
    return _sym(Co @ Co.T)


def _roothaan_fock(torch, Fa, Fb, Da, Db, S):
    """ROHF effective Fock matrix (asuka/hf/df_scf.py:691-719): with the core,
    open and virtual projectors pc = Db S, po = (Da - Db) S, pv = 1 - Da S and
    fc = (Fa + Fb)/2, F = G + G^T where G = 1/2 (pc^T fc pc + po^T fc po +
    pv^T fc pv) + po^T Fb pc + po^T Fa pv + pv^T fc pc."""
    nao = int(S.shape[0])
    fc = 0.5 * (Fa + Fb)
    pc = Db @ S
    po = (Da - Db) @ S
    pv = torch.eye(nao, dtype=torch.float64, device=S.device) - Da @ S
    G = 0.5 * (pc.T @ fc @ pc + po.T @ fc @ po + pv.T @ fc @ pv)
    G = G + po.T @ Fb @ pc + po.T @ Fa @ pv + pv.T @ fc @ pc
    return G + G.T


def _open_shell_scf(kind: str, S, hcore, jk_ctx, nalpha, nbeta, enuc, max_cycle, conv_tol, conv_tol_dm, diis,
                    diis_start_cycle, diis_space, damping, dm0, mo_coeff0, mixed_precision, profile,
                    mixed_until_dm=None):
    import torch

    dev = jk_ctx.device
    S = torch.as_tensor(np.asarray(S) if not isinstance(S, torch.Tensor) else S).to(dev, torch.float64)
    h = torch.as_tensor(np.asarray(hcore) if not isinstance(hcore, torch.Tensor) else hcore).to(dev, torch.float64)
    nao = int(S.shape[0])
    if tuple(S.shape) != (nao, nao) or tuple(h.shape) != (nao, nao):
        raise ValueError("S/hcore must be (nao, nao)")
    if nao != jk_ctx.nao:
        raise ValueError(f"nao mismatch: S has {nao}, DirectJKContext has {jk_ctx.nao}")
    na, nb = _occ_uhf(nalpha, nbeta, nao)
    if kind == "ROHF" and na < nb:
        raise ValueError("ROHF requires nalpha >= nbeta")
    occ_a = torch.zeros(nao, dtype=torch.float64, device=dev)
    occ_b = torch.zeros(nao, dtype=torch.float64, device=dev)
    occ_a[:na] = 1.0
    occ_b[:nb] = 1.0
    X = _orthogonalizer(torch, S)
    e0, C = _gen_eigh(torch, h, X)
    Ca = Cb = C
    ea = eb = e0

    def as_dev(A):
        return torch.as_tensor(np.asarray(A) if not isinstance(A, torch.Tensor) else A).to(dev, torch.float64)

    if dm0 is not None:
        if not isinstance(dm0, (tuple, list)) or len(dm0) != 2:
            raise TypeError(f"dm0 for {kind} must be a (Da, Db) tuple")
        Da, Db = _sym(as_dev(dm0[0])), _sym(as_dev(dm0[1]))
    else:
        if mo_coeff0 is not None:
            if isinstance(mo_coeff0, (tuple, list)):
                if kind == "ROHF" or len(mo_coeff0) != 2:
                    raise TypeError(f"mo_coeff0 for {kind} must be " + ("one matrix" if kind == "ROHF"
                                                                          else "a (Ca, Cb) tuple"))
                Ca, Cb = as_dev(mo_coeff0[0]), as_dev(mo_coeff0[1])
            else:
                Ca = Cb = as_dev(mo_coeff0)
        Da, Db = _density_spin(Ca, na), _density_spin(Cb, nb)

    lam = float(damping) if damping else 0.0
    nd = 2 if kind == "UHF" else 1
    diis_objs = [(_DIIS(torch, diis_space) if diis else None) for _ in range(nd)]
    e_last = None
    converged = False
    cycle = 0
    njk = 0
    mixed_now, mixed_until = _mixed_schedule(mixed_precision, mixed_until_dm)
    for cycle in range(1, int(max_cycle) + 1):
        Ja, Ka, Jb, Kb = direct_JK_multi(jk_ctx, Da, Db, mixed_precision=mixed_now)
        njk += 1
        J = Ja + Jb
        Fa = _sym(h + J - Ka)
        Fb = _sym(h + J - Kb)
        if kind == "UHF":
            Fs, Ds = [Fa, Fb], [Da, Db]
        else:
            Fs, Ds = [_roothaan_fock(torch, Fa, Fb, Da, Db, S)], [Da + Db]
        for k in range(nd):
            if diis_objs[k] is None or cycle < int(diis_start_cycle):
                continue
            M = Fs[k] @ Ds[k] @ S
            diis_objs[k].push(Fs[k], M - M.T)
            try:
                Ft = diis_objs[k].extrapolate()
            except Exception:
                Ft = None
            if Ft is None or not bool(torch.isfinite(Ft).all()):
                diis_objs[k] = _DIIS(torch, diis_space)
            else:
                Fs[k] = _sym(Ft)
        if kind == "UHF":
            ea, Ca = _gen_eigh(torch, Fs[0], X)
            eb, Cb = _gen_eigh(torch, Fs[1], X)
        else:
            ea, Ca = _gen_eigh(torch, Fs[0], X)
            eb, Cb = ea, Ca
        Da_n, Db_n = _density_spin(Ca, na), _density_spin(Cb, nb)
        if lam:
            Da_n = (1.0 - lam) * Da_n + lam * Da
            Db_n = (1.0 - lam) * Db_n + lam * Db
        Dt = Da_n + Db_n
        # the reference's energy expression (asuka/hf/direct_scf.py:543-548):
        # new densities with this cycle's J and K
        e_elec = float((Dt * h).sum()) + 0.5 * float((Dt * J).sum()) \
            - 0.5 * float((Da_n * Ka).sum()) - 0.5 * float((Db_n * Kb).sum())
        e_tot = e_elec + float(enuc)
        dm_err = float(torch.linalg.norm(Da_n - Da)) + float(torch.linalg.norm(Db_n - Db))
        de = abs(e_tot - e_last) if e_last is not None else float("inf")
        Da, Db, e_last = Da_n, Db_n, e_tot
        if de < conv_tol and dm_err < conv_tol_dm:
            converged = True
            break
        if mixed_now and dm_err < mixed_until:  # FP64 builds from here (see _MIXED_UNTIL_DM)
            mixed_now = False
    if profile is not None:
        profile.setdefault("scf", {})["iters"] = int(cycle)
        profile["scf"]["jk_calls"] = njk
    mo_e = (ea, eb) if kind == "UHF" else ea
    mo_c = (Ca, Cb) if kind == "UHF" else Ca
    return SCFResult(method=kind, converged=converged, niter=int(cycle),
                     e_tot=float(e_last if e_last is not None else float("nan")),
                     e_elec=float(e_last - float(enuc)) if e_last is not None else float("nan"),
                     e_nuc=float(enuc), mo_energy=mo_e, mo_coeff=mo_c, mo_occ=(occ_a, occ_b), dm=(Da, Db))


def uhf_direct(S, hcore, jk_ctx: DirectJKContext, *, nalpha: int, nbeta: int, enuc: float = 0.0,
               max_cycle: int = 50, conv_tol: float = 1e-10, conv_tol_dm: float = 1e-8, diis: bool = True,
               diis_start_cycle: int = 2, diis_space: int = 8, damping: float = 0.0, dm0=None, mo_coeff0=None,
               mixed_precision: bool | None = None, profile: dict | None = None,
               mixed_until_dm: float | None = None) -> SCFResult:
    """UHF with the two-density direct J/K build (asuka/hf/direct_scf.py:405-573):
    core guess, per-spin Fock h + J_a + J_b - K_s, one DIIS per spin on
    F D S - (F D S)^T, per-spin eigh, D_s = C_s,occ C_s,occ^T."""
    return _open_shell_scf("UHF", S, hcore, jk_ctx, nalpha, nbeta, enuc, max_cycle, conv_tol, conv_tol_dm, diis,
                           diis_start_cycle, diis_space, damping, dm0, mo_coeff0, mixed_precision, profile,
                           mixed_until_dm)


def rohf_direct(S, hcore, jk_ctx: DirectJKContext, *, nalpha: int, nbeta: int, enuc: float = 0.0,
                max_cycle: int = 50, conv_tol: float = 1e-10, conv_tol_dm: float = 1e-8, diis: bool = True,
                diis_start_cycle: int = 2, diis_space: int = 8, damping: float = 0.0, dm0=None, mo_coeff0=None,
                mixed_precision: bool | None = None, profile: dict | None = None,
                mixed_until_dm: float | None = None) -> SCFResult:
    """ROHF with the two-density direct J/K build (asuka/hf/direct_scf.py:576-):
    Roothaan's effective Fock matrix from the spin Fock matrices, one DIIS on
    F D_tot S - (F D_tot S)^T, one eigh, shared orbitals for both spins."""
    return _open_shell_scf("ROHF", S, hcore, jk_ctx, nalpha, nbeta, enuc, max_cycle, conv_tol, conv_tol_dm, diis,
                           diis_start_cycle, diis_space, damping, dm0, mo_coeff0, mixed_precision, profile,
                           mixed_until_dm)


__all__ = ["SCFResult", "rhf_direct", "rohf_direct", "uhf_direct"]
