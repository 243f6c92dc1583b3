"""ctypes binding of ``libasuka_jk.so`` (the C ABI declared in include/asuka_jk.h).

Loading fails loudly: there is no CPU fallback for any entry point. Error codes
map to the exception types of the reference's binding layer
(asuka/cueri/cuda/ext/src/cueri_cuda_ext_common.h:174-278):
ASUKA_E_INVALID -> ValueError, ASUKA_E_UNSUPPORTED -> NotImplementedError,
others -> RuntimeError.
"""

from __future__ import annotations

import ctypes
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
# ASUKA_JK_LIB points at another build of the same library (A/B timing of
# kernel variants); the in-tree build is the default
LIB_PATH = os.environ.get("ASUKA_JK_LIB") or os.path.join(_HERE, "libasuka_jk.so")

ASUKA_OK = 0
ASUKA_E_INVALID = -1
ASUKA_E_UNSUPPORTED = -2
ASUKA_E_HIP = -3
ASUKA_E_NOMEM = -4

WANT_J = 1
WANT_K = 2
MIXED = 4
COUNT = 8
TIME = 16
DETERMINISTIC = 32  # bitwise reproducible sums (64-bit fixed point, ASUKA_JK_DETERMINISTIC)

_i32p = ctypes.POINTER(ctypes.c_int32)
_i64p = ctypes.POINTER(ctypes.c_int64)
_f64p = ctypes.POINTER(ctypes.c_double)
_vp = ctypes.c_void_p


class BasisDesc(ctypes.Structure):
    _fields_ = [
        ("nshell", ctypes.c_int32),
        ("shell_cxyz", _f64p),
        ("shell_prim_start", _i32p),
        ("shell_nprim", _i32p),
        ("shell_l", _i32p),
        ("shell_ao_start", _i32p),
        ("shell_atom", _i32p),
        ("nprim", ctypes.c_int32),
        ("prim_exp", _f64p),
        ("prim_coef", _f64p),
        ("natm", ctypes.c_int32),
        ("atom_xyz", _f64p),
        ("atom_charge", _f64p),
    ]


class JKOpts(ctypes.Structure):
    _fields_ = [
        ("eps_schwarz", ctypes.c_double),
        ("eps_density", ctypes.c_double),
        ("tau_mixed", ctypes.c_double),
        ("streams", ctypes.c_int32),
        ("reserved", ctypes.c_int32 * 7),
    ]


class JKStats(ctypes.Structure):
    _fields_ = [
        ("class_pairs", ctypes.c_int64),
        ("schwarz_quartets", ctypes.c_int64),
        ("quartets", ctypes.c_int64),
        ("quartets_fp32", ctypes.c_int64),
        ("prim_quartets", ctypes.c_int64),
        ("prim_quartets_fp32", ctypes.c_int64),
        ("flops_efm", ctypes.c_double),
        ("flops_efm_fp32", ctypes.c_double),
        ("last_ms", ctypes.c_double),
        ("reserved", ctypes.c_int64 * 7),
    ]


# (name, restype, argtypes) for every symbol of include/asuka_jk.h
SYMBOLS = [
    ("asuka_jk_version", ctypes.c_int, []),
    ("asuka_jk_last_error", ctypes.c_char_p, []),
    ("asuka_jk_ctx_create", ctypes.c_int, [ctypes.POINTER(BasisDesc), ctypes.POINTER(JKOpts), _vp,
                                           ctypes.POINTER(_vp)]),
    ("asuka_jk_ctx_destroy", ctypes.c_int, [_vp]),
    ("asuka_jk_ctx_info", ctypes.c_int, [_vp, _i64p]),
    ("asuka_jk_build", ctypes.c_int, [_vp, _vp, _vp, _vp, ctypes.c_int, _vp]),
    ("asuka_jk_build_noalloc", ctypes.c_int, [_vp, _vp, _vp, _vp, ctypes.c_int, _vp]),
    ("asuka_scf_diis_solve", ctypes.c_int, [_vp, _vp, ctypes.c_int, _vp, _vp, _vp]),
    ("asuka_df_ctx_create", ctypes.c_int, [_vp, _vp, _vp, _vp]),
    ("asuka_df_ctx_destroy", ctypes.c_int, [_vp]),
    ("asuka_df_ctx_info", ctypes.c_int, [_vp, _vp]),
    ("asuka_df_metric", ctypes.c_int, [_vp, _vp, _vp]),
    ("asuka_df_int3c2e", ctypes.c_int, [_vp, _vp, _vp]),
    ("asuka_jk_build_multi", ctypes.c_int, [_vp, _vp, _vp, _vp, _vp, _vp, _vp, ctypes.c_int, _vp]),
    ("asuka_fock_build", ctypes.c_int, [_vp, _vp, _vp, _vp, ctypes.c_int, _vp]),
    ("asuka_jk_autotune", ctypes.c_int, [_vp, _vp, ctypes.c_int, ctypes.c_int, _vp]),
    ("asuka_jk_tuning_get", ctypes.c_int, [_vp, _i32p, ctypes.c_int]),
    ("asuka_jk_tuning_set", ctypes.c_int, [_vp, _i32p, ctypes.c_int, _vp]),
    ("asuka_jk_shell_pairs", ctypes.c_int, [_vp, _i32p, _i32p, _f64p]),
    ("asuka_jk_class_stats", ctypes.c_int, [_vp, _i64p, ctypes.c_int]),
    ("asuka_jk_last_timing", ctypes.c_int, [_vp, _f64p, _i64p]),
    ("asuka_jk_last_counts", ctypes.c_int, [_vp, _i64p, ctypes.c_int, _vp]),
    ("asuka_jk_stats", ctypes.c_int, [_vp, ctypes.POINTER(JKStats), _vp]),
    ("asuka_jk_eri_tiles", ctypes.c_int, [_vp, ctypes.c_int64, _vp, _vp, _vp, _vp, ctypes.c_int, _vp]),
    ("asuka_int1e", ctypes.c_int, [_vp, _vp, _vp, _vp, _vp]),
    ("asuka_jk_build_batched", ctypes.c_int, [ctypes.c_int, ctypes.POINTER(_vp), ctypes.POINTER(_vp),
                                              ctypes.POINTER(_vp), ctypes.POINTER(_vp), ctypes.c_int, _vp]),
    ("asuka_jk_batch_create", ctypes.c_int, [ctypes.c_int, ctypes.POINTER(ctypes.POINTER(BasisDesc)),
                                             ctypes.POINTER(JKOpts), _vp, ctypes.POINTER(_vp)]),
    ("asuka_jk_batch_info", ctypes.c_int, [_vp, _i32p, ctypes.c_int]),
    ("asuka_jk_batch_build", ctypes.c_int, [_vp, ctypes.POINTER(_vp), ctypes.POINTER(_vp), ctypes.POINTER(_vp),
                                            ctypes.c_int, _vp]),
    ("asuka_jk_grad", ctypes.c_int, [_vp, _vp, _vp, _vp, _vp, _vp]),
    ("asuka_int1e_grad", ctypes.c_int, [_vp, _vp, _vp, _vp, _vp]),
    ("asuka_rys_roots", ctypes.c_int, [ctypes.c_int, ctypes.c_double, _f64p, _f64p]),
    ("asuka_rys_roots_device", ctypes.c_int, [ctypes.c_int, ctypes.c_int64, _vp, _vp, _vp, _vp]),
]

_LIB = None


def lib() -> ctypes.CDLL:
    """Load the HIP library (build it first with ``asuka_amd._build.build()``)."""
    global _LIB
    if _LIB is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(
                f"{LIB_PATH} is missing: the HIP extension must be built (asuka_amd._build.build()); "
                "there is no CPU fallback")
        L = ctypes.CDLL(LIB_PATH)
        for name, res, args in SYMBOLS:
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _LIB = L
    return _LIB


def last_error() -> str:
    msg = lib().asuka_jk_last_error()
    return msg.decode() if msg else ""


def check(rc: int) -> int:
    if rc >= 0:
        return rc
    msg = last_error()
    if rc == ASUKA_E_INVALID:
        raise ValueError(msg)
    if rc == ASUKA_E_UNSUPPORTED:
        raise NotImplementedError(msg)
    raise RuntimeError(f"asuka_jk error {rc}: {msg}")


def ptr(t) -> int:
    """Device pointer of a torch tensor (0 for None)."""
    return 0 if t is None else int(t.data_ptr())


def stream_handle(device=None) -> int:
    import torch

    return int(torch.cuda.current_stream(device).cuda_stream)
