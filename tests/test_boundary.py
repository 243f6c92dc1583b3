"""The C-ABI library: loads without a GPU, exports every declared symbol, and its
host-side Rys table evaluation matches 40-digit golden values. No device calls."""

import ctypes
import os
import re

import numpy as np
import pytest

from asuka_amd import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared_symbols():
    text = open(os.path.join(ROOT, "include", "asuka_jk.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(asuka_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    L = _lib.lib()
    declared = _declared_symbols()
    assert len(declared) >= 14
    for name in declared:
        assert hasattr(L, name), name
    bound = {n for n, _, _ in _lib.SYMBOLS}
    assert set(declared) == bound


def test_struct_layouts():
    import ctypes

    # asuka_jk_opts: 3 doubles + int32 streams + int32[7]; asuka_jk_stats_out: 6 int64 + 3 doubles + int64[7]
    assert ctypes.sizeof(_lib.JKOpts) == 56
    assert ctypes.sizeof(_lib.JKStats) == 128


def test_version_and_error_paths():
    L = _lib.lib()
    assert L.asuka_jk_version() == 1
    x = np.zeros(8)
    w = np.zeros(8)
    with pytest.raises(ValueError):
        _lib.check(L.asuka_rys_roots(9, 1.0, x.ctypes.data_as(_lib._f64p), w.ctypes.data_as(_lib._f64p)))
    with pytest.raises(ValueError):
        _lib.check(L.asuka_rys_roots(2, -1.0, x.ctypes.data_as(_lib._f64p), w.ctypes.data_as(_lib._f64p)))
    with pytest.raises(ValueError):
        _lib.check(L.asuka_jk_ctx_create(None, None, None, None))


@pytest.mark.parametrize("n", range(1, 8))
def test_host_table_roots_against_golden(golden, n):
    L = _lib.lib()
    for i, T in enumerate(golden["T"]):
        x = np.zeros(n)
        w = np.zeros(n)
        _lib.check(L.asuka_rys_roots(n, float(T), x.ctypes.data_as(_lib._f64p), w.ctypes.data_as(_lib._f64p)))
        np.testing.assert_allclose(x, golden["x"][n, i, :n], rtol=5e-15, atol=0)
        np.testing.assert_allclose(w, golden["w"][n, i, :n], rtol=5e-15, atol=0)


def test_host_table_moments_dense_scan(golden):
    """Between the golden points: sum_u w_u x_u^k = F_k(T) against the
    independent Boys series, on a dense T grid crossing every interval edge and
    the asymptotic switch."""
    import mmd_ref

    L = _lib.lib()
    for n in (1, 3, 5, 7):
        for T in np.concatenate([np.linspace(0, 100, 1001), [1e-300, 37.999999, 38.0, 75.999, 76.0]]):
            x = np.zeros(n)
            w = np.zeros(n)
            _lib.check(L.asuka_rys_roots(n, float(T), x.ctypes.data_as(_lib._f64p), w.ctypes.data_as(_lib._f64p)))
            F = mmd_ref.boys_array(2 * n - 1, float(T))
            for k in range(2 * n):
                assert abs(np.sum(w * x ** k) - F[k]) <= 4e-15 * F[0], (n, T, k)


def test_scf_mixed_precision_schedule(monkeypatch):
    """The SCF drivers' precision schedule: FP64 when the build is not mixed;
    mixed builds switch to FP64 below ||dD||_F = 1e-3 by default, overridden
    by ASUKA_SCF_MIXED_UNTIL_DM or the argument; 0 keeps them mixed."""
    from asuka_amd.hf.direct_scf import _mixed_schedule

    monkeypatch.delenv("ASUKA_SCF_MIXED_UNTIL_DM", raising=False)
    assert _mixed_schedule(False, None) == (False, 0.0)
    assert _mixed_schedule(False, 1e-2) == (False, 0.0)
    assert _mixed_schedule(True, None) == (True, 1e-3)
    assert _mixed_schedule(True, 5e-4) == (True, 5e-4)
    assert _mixed_schedule(True, 0.0) == (True, 0.0)
    monkeypatch.setenv("ASUKA_SCF_MIXED_UNTIL_DM", "2e-3")
    assert _mixed_schedule(True, None) == (True, 2e-3)
    assert _mixed_schedule(True, 1e-4) == (True, 1e-4)
    monkeypatch.setenv("ASUKA_ERI_MIXED_PRECISION", "0")
    assert _mixed_schedule(None, None) == (False, 0.0)
    monkeypatch.setenv("ASUKA_ERI_MIXED_PRECISION", "1")
    assert _mixed_schedule(None, None) == (True, 2e-3)
