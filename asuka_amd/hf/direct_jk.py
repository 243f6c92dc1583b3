"""Integral-direct J/K and Fock builds on the MI355X engine.

Drop-in mirror of the reference operator API ``asuka.hf.direct_jk``
(asuka/hf/direct_jk.py:3098-3106 ``__all__``):

* ``make_direct_jk_context(ao_basis, *, eps_schwarz, threads, max_tile_bytes,
  max_slab_tasks, gpu_task_budget_bytes)`` (:1143) -> :class:`DirectJKContext`
* ``direct_JK(ctx, D, *, want_J, want_K, profile, stats, tile_dtype,
  mixed_precision, f32_accum)`` (:1530) -> ``(J, K)``
* ``direct_fock_rhf(ctx, D, hcore, ...)`` (:2205) -> ``F``
* ``direct_JK_multi(ctx, Da, Db, ...)`` (:2716) -> ``(Ja, Ka, Jb, Kb)``

Arrays are torch tensors on the GPU (float64, row-major (nao, nao)); numpy
inputs are copied to the device. J, K and F are returned symmetrised. Shape
errors raise ``ValueError`` with "must be" in the message, as the reference
(:1591-1592); angular momentum beyond f raises ``NotImplementedError``
(:1197-1200). All work runs in ``libasuka_jk.so`` through its C ABI; there is
no CPU fallback.

The reference's ``threads``, ``max_tile_bytes``, ``max_slab_tasks`` and
``gpu_task_budget_bytes`` size its CUDA launches and task slabs; this engine
generates quartets on the device from per-class row lists and has no slabs or
tile buffers, so those arguments are accepted and ignored.
"""

from __future__ import annotations

import ctypes
import os
import time
from dataclasses import dataclass, field
from typing import Any

import numpy as np

from .. import _lib


def _torch():
    import torch

    if not torch.cuda.is_available():
        raise RuntimeError("the direct J/K engine needs a GPU (torch.cuda.is_available() is False)")
    return torch


@dataclass
class DirectJKContext:
    """Reusable screening/sorting state of one basis on one device
    (asuka/hf/direct_jk.py:729-760)."""

    ao_basis: Any
    nao: int
    max_l: int
    nsp: int
    ntasks: int
    eps_schwarz: float
    eps_density: float
    device: Any
    handle: int
    sp_A: np.ndarray
    sp_B: np.ndarray
    Q_sp: np.ndarray
    natm: int = 0
    nprim_pairs: int = 0
    atom_coords_bohr: np.ndarray | None = field(default=None, repr=False)
    _keep: tuple = field(default=(), repr=False)

    def __del__(self):
        h = getattr(self, "handle", 0)
        if h:
            try:
                _lib.lib().asuka_jk_ctx_destroy(ctypes.c_void_p(h))
            except Exception:
                pass
            self.handle = 0

    def class_stats(self) -> list[dict]:
        """Per work segment (class pair and ket contraction depth): (la, lb, lc,
        ld), Schwarz-screened shell quartets, primitive quartets and
        primitive-quartet roots, and the primitive quartets counted before
        primitive-pair screening (prim_quartets0)."""
        info = np.zeros(8, dtype=np.int64)
        _lib.check(_lib.lib().asuka_jk_ctx_info(ctypes.c_void_p(self.handle), info.ctypes.data_as(_lib._i64p)))
        cap = max(1, int(info[5]))
        rows = np.zeros((cap, 8), dtype=np.int64)
        n = _lib.check(_lib.lib().asuka_jk_class_stats(ctypes.c_void_p(self.handle),
                                                       rows.ctypes.data_as(_lib._i64p), cap))
        out = []
        for r in rows[:n]:
            out.append({"class": tuple(int(x) for x in r[:4]), "quartets": int(r[4]),
                        "prim_quartets": int(r[5]), "prim_roots": int(r[6]), "prim_quartets0": int(r[7])})
        return out

    def last_counts(self) -> list[dict]:
        """Counters of the last build run with the count flag (``stats=`` of
        :func:`direct_JK`): per class pair, in :meth:`class_stats` order, the
        quartets evaluated after density screening by symmetry bin
        (A!=B)<<2 | (C!=D)<<1 | (bra pair != ket pair), their primitive
        quartets, the FP32 share of both, and for the class-specialised kernels
        the primitive quartets their loops evaluate and the lane slots those
        loops take (their ratio is the loops' SIMD utilisation)."""
        cls = self.class_stats()
        n = len(cls)
        rows = np.zeros((max(1, n), 16), dtype=np.int64)
        _lib.check(_lib.lib().asuka_jk_last_counts(ctypes.c_void_p(self.handle), rows.ctypes.data_as(_lib._i64p), n,
                                                   ctypes.c_void_p(_lib.stream_handle(self.device))))
        return [{"class": c["class"], "bins": rows[i, :8].copy(), "prim_quartets": int(rows[i, 8]),
                 "quartets_fp32": int(rows[i, 9]), "prim_quartets_fp32": int(rows[i, 10]),
                 "prim_quartets_evaluated": int(rows[i, 11]), "prim_lane_slots": int(rows[i, 12])}
                for i, c in enumerate(cls)]

    def stats(self) -> dict:
        """Work summary (asuka_jk_stats): Schwarz-screened quartets and, from the
        last build with the count flag, the quartets and primitive quartets
        evaluated, their FP32 share and the EFM-v1 FLOPs."""
        out = _lib.JKStats()
        _lib.check(_lib.lib().asuka_jk_stats(ctypes.c_void_p(self.handle), ctypes.byref(out),
                                             ctypes.c_void_p(_lib.stream_handle(self.device))))
        return {k: getattr(out, k) for k, _ in _lib.JKStats._fields_ if k != "reserved"}

    def autotune(self, D, *, mixed_precision: bool | None = None, reps: int = 2) -> list[dict]:
        """Time every class pair's launch on density D and keep the fastest
        launch shape per class pair (see :meth:`tuning`); the class pairs are
        then spread over the context's streams by their measured times.
        Synchronous. Counterpart of the reference's online hot-class autotune
        (asuka/cueri/gpu.py:4929-5005)."""
        torch = _torch()
        Dd = _as_device_matrix(torch, D, self.nao, "D", self.device)
        _lib.check(_lib.lib().asuka_jk_autotune(ctypes.c_void_p(self.handle), _lib.ptr(Dd), _flags(mixed_precision),
                                                int(reps), ctypes.c_void_p(_lib.stream_handle(self.device))))
        return self.tuning()

    def tuning(self) -> list[dict]:
        """The tuning table: per class pair (la, lb, lc, ld), wavefronts per
        workgroup, wave units per row task, LDS row accumulators on/off and the
        tuned launch time in ms (0.0 if not tuned)."""
        cap = 4096
        rows = np.zeros((cap, 8), dtype=np.int32)
        n = _lib.check(_lib.lib().asuka_jk_tuning_get(ctypes.c_void_p(self.handle),
                                                      rows.ctypes.data_as(_lib._i32p), cap))
        return [{"class": tuple(int(x) for x in r[:4]), "waves_per_group": int(r[4]), "units_per_task": int(r[5]),
                 "lds_accumulators": bool(r[6]), "tuned_ms": int(r[7]) / 1000.0} for r in rows[:min(n, cap)]]

    def set_tuning(self, table: list[dict]) -> None:
        """Apply a tuning table (from :meth:`autotune` or :meth:`tuning`, possibly
        of another molecule in the same basis); classes absent here are skipped."""
        rows = np.array([list(t["class"]) + [int(t["waves_per_group"]), int(t["units_per_task"]),
                                              int(bool(t["lds_accumulators"])), int(round(1000 * t["tuned_ms"]))]
                         for t in table], dtype=np.int32).reshape(-1, 8)
        rows = np.ascontiguousarray(rows)
        _lib.check(_lib.lib().asuka_jk_tuning_set(ctypes.c_void_p(self.handle), rows.ctypes.data_as(_lib._i32p),
                                                  int(rows.shape[0]), ctypes.c_void_p(_lib.stream_handle(self.device))))


def _basis_desc(ao_basis, atom_xyz=None, atom_charge=None):
    keep = []

    def arr(x, dt):
        a = np.ascontiguousarray(np.asarray(x, dtype=dt))
        keep.append(a)
        return a

    cxyz = arr(ao_basis.shell_cxyz, np.float64)
    ps = arr(ao_basis.shell_prim_start, np.int32)
    npr = arr(ao_basis.shell_nprim, np.int32)
    sl = arr(ao_basis.shell_l, np.int32)
    ao = arr(ao_basis.shell_ao_start, np.int32)
    satom = getattr(ao_basis, "shell_atom", None)
    sat = arr(satom, np.int32) if satom is not None else None
    pe = arr(ao_basis.prim_exp, np.float64)
    pc = arr(ao_basis.prim_coef, np.float64)
    natm = 0
    axyz = az = None
    if atom_xyz is not None:
        axyz = arr(np.asarray(atom_xyz).reshape(-1, 3), np.float64)
        az = arr(atom_charge, np.float64)
        natm = int(axyz.shape[0])
    d = _lib.BasisDesc(
        int(sl.shape[0]), cxyz.ctypes.data_as(_lib._f64p), ps.ctypes.data_as(_lib._i32p),
        npr.ctypes.data_as(_lib._i32p), sl.ctypes.data_as(_lib._i32p), ao.ctypes.data_as(_lib._i32p),
        sat.ctypes.data_as(_lib._i32p) if sat is not None else None, int(pe.shape[0]),
        pe.ctypes.data_as(_lib._f64p), pc.ctypes.data_as(_lib._f64p), natm,
        axyz.ctypes.data_as(_lib._f64p) if axyz is not None else None,
        az.ctypes.data_as(_lib._f64p) if az is not None else None)
    return d, keep


def make_direct_jk_context(
    ao_basis,
    *,
    eps_schwarz: float = 1e-12,
    threads: int = 256,
    max_tile_bytes: int = 256 << 20,
    max_slab_tasks: int | None = 200_000_000,
    gpu_task_budget_bytes: int | None = 2 << 30,
    eps_density: float | None = None,
    tau_mixed: float = 0.0,
    atom_coords_bohr=None,
    atom_charges=None,
    streams: int | None = None,
) -> DirectJKContext:
    """One-time setup (asuka/hf/direct_jk.py:1143-1470): shell pairs,
    primitive-pair tables, Schwarz bounds on the GPU, Q-sorted class work lists.

    ``eps_density`` defaults to ``$ASUKA_DIRECT_JK_EPS_DENSITY`` or
    max(eps_schwarz, 1e-10) (:1594-1596). Atom coordinates/charges are needed
    only for the one-electron integrals. ``streams`` (default
    ``$ASUKA_DIRECT_JK_STREAMS`` or 4) is the number of streams a build's
    class-pair launches are spread over, the caller's included.
    """
    torch = _torch()
    del threads, max_tile_bytes, max_slab_tasks, gpu_task_budget_bytes
    if eps_density is None:
        env = os.environ.get("ASUKA_DIRECT_JK_EPS_DENSITY", "").strip()
        eps_density = float(env) if env else max(float(eps_schwarz), 1e-10)
    if int(np.asarray(ao_basis.shell_l).size) == 0:
        raise ValueError("AO basis has zero shells")
    desc, keep = _basis_desc(ao_basis, atom_coords_bohr, atom_charges)
    if streams is None:
        env = os.environ.get("ASUKA_DIRECT_JK_STREAMS", "").strip()
        streams = int(env) if env else 0
    opts = _lib.JKOpts(float(eps_schwarz), float(eps_density), float(tau_mixed), int(streams))
    device = torch.device("cuda", torch.cuda.current_device())
    h = ctypes.c_void_p()
    _lib.check(_lib.lib().asuka_jk_ctx_create(ctypes.byref(desc), ctypes.byref(opts),
                                              ctypes.c_void_p(_lib.stream_handle()), ctypes.byref(h)))
    info = np.zeros(8, dtype=np.int64)
    _lib.check(_lib.lib().asuka_jk_ctx_info(h, info.ctypes.data_as(_lib._i64p)))
    nsp = int(info[2])
    sp_A = np.zeros(nsp, np.int32)
    sp_B = np.zeros(nsp, np.int32)
    Q = np.zeros(nsp, np.float64)
    _lib.check(_lib.lib().asuka_jk_shell_pairs(h, sp_A.ctypes.data_as(_lib._i32p), sp_B.ctypes.data_as(_lib._i32p),
                                               Q.ctypes.data_as(_lib._f64p)))
    return DirectJKContext(ao_basis=ao_basis, nao=int(info[0]), max_l=int(info[4]), nsp=nsp, ntasks=int(info[3]),
                           eps_schwarz=float(eps_schwarz), eps_density=float(eps_density), device=device,
                           handle=int(h.value), sp_A=sp_A, sp_B=sp_B, Q_sp=Q, natm=int(info[6]),
                           nprim_pairs=int(info[7]),
                           atom_coords_bohr=(None if atom_coords_bohr is None
                                             else np.array(atom_coords_bohr, dtype=np.float64).reshape(-1, 3)),
                           _keep=tuple(keep))


def _as_device_matrix(torch, X, nao: int, name: str, device):
    if not isinstance(X, torch.Tensor):
        X = torch.as_tensor(np.asarray(X, dtype=np.float64))
    if X.ndim != 2 or tuple(X.shape) != (nao, nao):
        raise ValueError(f"{name} must be ({nao}, {nao}), got {tuple(X.shape)}")
    return X.to(device=device, dtype=torch.float64).contiguous()


def _flags(mixed_precision: bool | None) -> int:
    """Precision flag of a build. None takes the reference's default, mixed
    precision unless ``ASUKA_ERI_MIXED_PRECISION=0`` (asuka/hf/direct_jk.py:
    1580-1581)."""
    if mixed_precision is None:
        mixed_precision = os.environ.get("ASUKA_ERI_MIXED_PRECISION", "1") != "0"
    return _lib.MIXED if mixed_precision else 0


def direct_JK(
    ctx: DirectJKContext,
    D,
    *,
    want_J: bool = True,
    want_K: bool = True,
    profile: dict | None = None,
    stats: dict | None = None,
    tile_dtype: str | None = None,
    mixed_precision: bool | None = None,
    f32_accum: bool | None = None,
    out: tuple | None = None,
    deterministic: bool = False,
):
    """J and K for density D (asuka/hf/direct_jk.py:1530-2202).

    Returns ``(J, K)`` as symmetrised (nao, nao) float64 GPU tensors (or None
    for an output not wanted). ``out=(J, K)`` reuses caller buffers.
    ``tile_dtype`` and ``f32_accum`` are accepted for signature compatibility;
    this engine never materialises tiles and accumulates in FP64.
    ``deterministic``: bitwise reproducible J and K (ASUKA_JK_DETERMINISTIC:
    contributions rounded to 2^-44 and summed as 64-bit integers).
    """
    del tile_dtype, f32_accum
    if not want_J and not want_K:
        return None, None
    torch = _torch()
    nao = ctx.nao
    Dd = _as_device_matrix(torch, D, nao, "D", ctx.device)
    if out is not None:
        J, K = out
    else:
        J = torch.empty((nao, nao), dtype=torch.float64, device=ctx.device) if want_J else None
        K = torch.empty((nao, nao), dtype=torch.float64, device=ctx.device) if want_K else None
    flags = (_lib.WANT_J if want_J else 0) | (_lib.WANT_K if want_K else 0) | _flags(mixed_precision)
    if stats is not None:
        flags |= _lib.COUNT
    if deterministic:
        flags |= _lib.DETERMINISTIC
    t0 = time.perf_counter()
    _lib.check(_lib.lib().asuka_jk_build(ctypes.c_void_p(ctx.handle), _lib.ptr(Dd), _lib.ptr(J if want_J else None),
                                         _lib.ptr(K if want_K else None), flags,
                                         ctypes.c_void_p(_lib.stream_handle(ctx.device))))
    if profile is not None:
        torch.cuda.synchronize(ctx.device)
        profile["direct_jk_t_s"] = float(time.perf_counter() - t0)
        profile["direct_jk_ntasks"] = int(ctx.ntasks)
    if stats is not None:
        _fill_stats(ctx, stats)
    return (J if want_J else None), (K if want_K else None)


def _fill_stats(ctx: DirectJKContext, stats: dict) -> None:
    """The stats dict of a counted build (asuka/hf/direct_jk.py:1687-1700):
    Schwarz-screened and evaluated quartets, per class pair."""
    counts = ctx.last_counts()
    stats["direct_jk_ntasks"] = int(ctx.ntasks)
    stats["classes"] = ctx.class_stats()
    stats["class_counts"] = counts
    stats["direct_jk_evaluated"] = int(sum(int(c["bins"].sum()) for c in counts))


def direct_JK_multi(
    ctx: DirectJKContext,
    Da,
    Db,
    *,
    want_J: bool = True,
    want_K: bool = True,
    profile: dict | None = None,
    stats: dict | None = None,
    tile_dtype: str | None = None,
    mixed_precision: bool | None = None,
    f32_accum: bool | None = None,
    out: tuple | None = None,
    deterministic: bool = False,
):
    """(Ja, Ka, Jb, Kb) for two densities, each ERI evaluated once
    (asuka/hf/direct_jk.py:2716-3095).

    The UHF/ROHF build: every screened quartet is digested with Da into
    (Ja, Ka) and with Db into (Jb, Kb) in the same kernel, so the cost is one
    ERI pass instead of two. Returns symmetrised (nao, nao) float64 GPU
    tensors (None for outputs not wanted); ``out=(Ja, Ka, Jb, Kb)`` reuses
    caller buffers. ``deterministic``: bitwise reproducible outputs (see
    :func:`direct_JK`).
    """
    del tile_dtype, f32_accum
    if not want_J and not want_K:
        return None, None, None, None
    torch = _torch()
    nao = ctx.nao
    Dad = _as_device_matrix(torch, Da, nao, "Da", ctx.device)
    Dbd = _as_device_matrix(torch, Db, nao, "Db", ctx.device)

    def _new():
        return torch.empty((nao, nao), dtype=torch.float64, device=ctx.device)

    if out is not None:
        Ja, Ka, Jb, Kb = out
    else:
        Ja, Jb = (_new(), _new()) if want_J else (None, None)
        Ka, Kb = (_new(), _new()) if want_K else (None, None)
    if not want_J:
        Ja = Jb = None
    if not want_K:
        Ka = Kb = None
    flags = (_lib.WANT_J if want_J else 0) | (_lib.WANT_K if want_K else 0) | _flags(mixed_precision)
    if deterministic:
        flags |= _lib.DETERMINISTIC
    t0 = time.perf_counter()
    _lib.check(_lib.lib().asuka_jk_build_multi(ctypes.c_void_p(ctx.handle), _lib.ptr(Dad), _lib.ptr(Dbd),
                                               _lib.ptr(Ja), _lib.ptr(Ka), _lib.ptr(Jb), _lib.ptr(Kb), flags,
                                               ctypes.c_void_p(_lib.stream_handle(ctx.device))))
    if profile is not None:
        torch.cuda.synchronize(ctx.device)
        profile["direct_jk_t_s"] = float(time.perf_counter() - t0)
        profile["direct_jk_ntasks"] = int(ctx.ntasks)
    if stats is not None:
        stats["direct_jk_ntasks"] = int(ctx.ntasks)
        stats["classes"] = ctx.class_stats()
    return Ja, Ka, Jb, Kb


def direct_fock_rhf(
    ctx: DirectJKContext,
    D,
    hcore,
    *,
    profile: dict | None = None,
    stats: dict | None = None,
    tile_dtype: str | None = None,
    mixed_precision: bool | None = None,
    f32_accum: bool | None = None,
    out=None,
    deterministic: bool = False,
):
    """F = h + J(D) - K(D)/2, symmetrised (asuka/hf/direct_jk.py:2205-2713).
    ``deterministic``: bitwise reproducible F (see :func:`direct_JK`)."""
    del tile_dtype, f32_accum
    torch = _torch()
    nao = ctx.nao
    Dd = _as_device_matrix(torch, D, nao, "D", ctx.device)
    hd = _as_device_matrix(torch, hcore, nao, "hcore", ctx.device)
    F = out if out is not None else torch.empty((nao, nao), dtype=torch.float64, device=ctx.device)
    t0 = time.perf_counter()
    flags = _flags(mixed_precision) | (_lib.COUNT if stats is not None else 0)
    if deterministic:
        flags |= _lib.DETERMINISTIC
    _lib.check(_lib.lib().asuka_fock_build(ctypes.c_void_p(ctx.handle), _lib.ptr(Dd), _lib.ptr(hd), _lib.ptr(F),
                                           flags, ctypes.c_void_p(_lib.stream_handle(ctx.device))))
    if profile is not None:
        torch.cuda.synchronize(ctx.device)
        profile["direct_fock_t_s"] = float(time.perf_counter() - t0)
        profile["direct_fock_ntasks"] = int(ctx.ntasks)
    if stats is not None:
        _fill_stats(ctx, stats)
    return F


@dataclass
class DirectJKBatchContext:
    """Several independent molecules in one context (asuka_jk_batch_create):
    one launch per class pair covers the whole batch."""

    naos: list
    device: Any
    handle: int
    ntasks: int
    _keep: tuple = field(default=(), repr=False)

    def __del__(self):
        h = getattr(self, "handle", 0)
        if h:
            try:
                _lib.lib().asuka_jk_ctx_destroy(ctypes.c_void_p(h))
            except Exception:
                pass
            self.handle = 0

    def set_tuning(self, table: list[dict]) -> None:
        """Apply a tuning table (e.g. from one molecule's
        :meth:`DirectJKContext.autotune` in the same basis)."""
        DirectJKContext.set_tuning(self, table)

    def tuning(self) -> list[dict]:
        return DirectJKContext.tuning(self)


def make_direct_jk_batch_context(ao_bases, *, eps_schwarz: float = 1e-12, eps_density: float | None = None,
                                 tau_mixed: float = 0.0, streams: int | None = None) -> DirectJKBatchContext:
    """One context for a batch of independent molecules (one AO basis each)."""
    torch = _torch()
    if eps_density is None:
        env = os.environ.get("ASUKA_DIRECT_JK_EPS_DENSITY", "").strip()
        eps_density = float(env) if env else max(float(eps_schwarz), 1e-10)
    if streams is None:
        env = os.environ.get("ASUKA_DIRECT_JK_STREAMS", "").strip()
        streams = int(env) if env else 0
    ao_bases = list(ao_bases)
    if not ao_bases:
        raise ValueError("ao_bases must hold at least one basis")
    descs, keeps = [], []
    for ao in ao_bases:
        if int(np.asarray(ao.shell_l).size) == 0:
            raise ValueError("AO basis has zero shells")
        d, k = _basis_desc(ao)
        descs.append(d)
        keeps.append(k)
    arr = (ctypes.POINTER(_lib.BasisDesc) * len(descs))(*[ctypes.pointer(d) for d in descs])
    opts = _lib.JKOpts(float(eps_schwarz), float(eps_density), float(tau_mixed), int(streams))
    device = torch.device("cuda", torch.cuda.current_device())
    h = ctypes.c_void_p()
    _lib.check(_lib.lib().asuka_jk_batch_create(len(descs), arr, ctypes.byref(opts),
                                                ctypes.c_void_p(_lib.stream_handle()), ctypes.byref(h)))
    naos = np.zeros(len(descs), dtype=np.int32)
    _lib.check(_lib.lib().asuka_jk_batch_info(h, naos.ctypes.data_as(_lib._i32p), len(descs)))
    info = np.zeros(8, dtype=np.int64)
    _lib.check(_lib.lib().asuka_jk_ctx_info(h, info.ctypes.data_as(_lib._i64p)))
    return DirectJKBatchContext(naos=[int(n) for n in naos], device=device, handle=int(h.value), ntasks=int(info[3]),
                                _keep=(descs, keeps))


def direct_JK_batched(ctxs, Ds, *, want_J: bool = True, want_K: bool = True, mixed_precision: bool | None = None,
                      out=None):
    """J and K of several independent molecules in one call, as a list of
    (J, K). ``ctxs`` is either a :class:`DirectJKBatchContext` (one launch per
    class pair for the whole batch, asuka_jk_batch_build) or a list of
    per-molecule contexts (their builds issued in turn, asuka_jk_build_batched).
    No host synchronisation either way, so a HIP graph can capture the batch."""
    torch = _torch()
    if isinstance(ctxs, DirectJKBatchContext):
        bc = ctxs
        n = len(bc.naos)
        if len(Ds) != n:
            raise ValueError("Ds must have one density per molecule")
        Dd = [_as_device_matrix(torch, D, m, "D", bc.device) for m, D in zip(bc.naos, Ds)]
        if out is None:
            out = [(torch.empty((m, m), dtype=torch.float64, device=bc.device) if want_J else None,
                    torch.empty((m, m), dtype=torch.float64, device=bc.device) if want_K else None) for m in bc.naos]
        arr = ctypes.c_void_p * n
        dp = arr(*[ctypes.c_void_p(_lib.ptr(d)) for d in Dd])
        jp = arr(*[ctypes.c_void_p(_lib.ptr(o[0])) for o in out])
        kp = arr(*[ctypes.c_void_p(_lib.ptr(o[1])) for o in out])
        flags = (_lib.WANT_J if want_J else 0) | (_lib.WANT_K if want_K else 0) | _flags(mixed_precision)
        _lib.check(_lib.lib().asuka_jk_batch_build(ctypes.c_void_p(bc.handle), dp, jp, kp, flags,
                                                   ctypes.c_void_p(_lib.stream_handle(bc.device))))
        return out
    n = len(ctxs)
    if len(Ds) != n:
        raise ValueError("Ds must have one density per context")
    Dd = [_as_device_matrix(torch, D, c.nao, "D", c.device) for c, D in zip(ctxs, Ds)]
    if out is None:
        out = [(torch.empty((c.nao, c.nao), dtype=torch.float64, device=c.device) if want_J else None,
                torch.empty((c.nao, c.nao), dtype=torch.float64, device=c.device) if want_K else None) for c in ctxs]
    arr = ctypes.c_void_p * max(1, n)
    hs = arr(*[ctypes.c_void_p(c.handle) for c in ctxs])
    dp = arr(*[ctypes.c_void_p(_lib.ptr(d)) for d in Dd])
    jp = arr(*[ctypes.c_void_p(_lib.ptr(o[0])) for o in out])
    kp = arr(*[ctypes.c_void_p(_lib.ptr(o[1])) for o in out])
    flags = (_lib.WANT_J if want_J else 0) | (_lib.WANT_K if want_K else 0) | _flags(mixed_precision)
    _lib.check(_lib.lib().asuka_jk_build_batched(n, hs, dp, jp, kp, flags,
                                                 ctypes.c_void_p(_lib.stream_handle(ctxs[0].device if n else None))))
    return out


def eri_tiles(ctx: DirectJKContext, spab, spcd):
    """Contracted ERI tiles for explicit (spAB, spCD) quartets (test hook).
    Returns a list of (nAB, nCD) numpy arrays in shell-pair orientation."""
    torch = _torch()
    L = np.asarray(ctx.ao_basis.shell_l)
    nc = lambda l: (l + 1) * (l + 2) // 2
    spab = np.asarray(spab, dtype=np.int32)
    spcd = np.asarray(spcd, dtype=np.int32)
    shapes = [(nc(L[ctx.sp_A[a]]) * nc(L[ctx.sp_B[a]]), nc(L[ctx.sp_A[c]]) * nc(L[ctx.sp_B[c]]))
              for a, c in zip(spab, spcd)]
    sizes = np.array([r * c for r, c in shapes], dtype=np.int64)
    off = np.zeros(len(sizes), dtype=np.int64)
    if len(sizes) > 1:
        off[1:] = np.cumsum(sizes)[:-1]
    dev = ctx.device
    t_ab = torch.as_tensor(spab, device=dev)
    t_cd = torch.as_tensor(spcd, device=dev)
    t_off = torch.as_tensor(off, device=dev)
    buf = torch.zeros(int(sizes.sum()) if len(sizes) else 1, dtype=torch.float64, device=dev)
    _lib.check(_lib.lib().asuka_jk_eri_tiles(ctypes.c_void_p(ctx.handle), int(len(spab)), _lib.ptr(t_ab),
                                             _lib.ptr(t_cd), _lib.ptr(t_off), _lib.ptr(buf), 0,
                                             ctypes.c_void_p(_lib.stream_handle(dev))))
    host = buf.cpu().numpy()
    return [host[o:o + s].reshape(shp) for o, s, shp in zip(off, sizes, shapes)]


__all__ = ["DirectJKBatchContext", "DirectJKContext", "direct_JK", "direct_JK_batched", "direct_JK_multi",
           "direct_fock_rhf", "eri_tiles", "make_direct_jk_batch_context", "make_direct_jk_context"]
