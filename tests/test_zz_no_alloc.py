"""asuka_jk_build_noalloc (per-build scratch in a static device array, no
allocation) against asuka_jk_build, through the C ABI: two calls in a row on
two contexts, the second one smaller, in FP64 and mixed precision."""

import ctypes

import numpy as np
import pytest

from asuka_amd.basis import build_ao_basis_cart
from asuka_amd.molecule import water_cluster

pytestmark = pytest.mark.gpu


def _density(nao, nocc, seed):
    C = np.random.default_rng(seed).standard_normal((nao, nocc))
    C, _ = np.linalg.qr(C)
    return 2.0 * C @ C.T


def test_noalloc_build_matches_build(gpu):
    from asuka_amd import _lib
    from asuka_amd.hf.direct_jk import make_direct_jk_context

    torch = gpu
    L = _lib.lib()
    sptr = ctypes.c_void_p(_lib.stream_handle())
    for n, basis in ((4, "cc-pVTZ"), (2, "cc-pVDZ")):  # the second call is smaller
        print(f"\n[noalloc] (H2O){n} {basis}", flush=True)
        ao, _ = build_ao_basis_cart(water_cluster(n, n, basis))
        ctx = make_direct_jk_context(ao, eps_schwarz=1e-12)
        h = ctypes.c_void_p(ctx.handle)
        D = torch.as_tensor(_density(ao.nao, 5 * n, 40 + n), device="cuda")
        for mixed in (False, True):
            flags = _lib.WANT_J | _lib.WANT_K | (_lib.MIXED if mixed else 0)
            J0, K0 = torch.empty_like(D), torch.empty_like(D)
            J1, K1 = torch.full_like(D, 5.0), torch.full_like(D, -5.0)
            _lib.check(L.asuka_jk_build(h, _lib.ptr(D), _lib.ptr(J0), _lib.ptr(K0), flags, sptr))
            _lib.check(L.asuka_jk_build_noalloc(h, _lib.ptr(D), _lib.ptr(J1), _lib.ptr(K1), flags, sptr))
            torch.cuda.synchronize()
            scale = max(1.0, float(J0.abs().max()), float(K0.abs().max()))
            assert float((J1 - J0).abs().max()) < 1e-12 * scale, (n, mixed)
            assert float((K1 - K0).abs().max()) < 1e-12 * scale, (n, mixed)
            assert float((J1 - J1.T).abs().max()) == 0.0
        with pytest.raises(ValueError, match="TIME"):
            _lib.check(L.asuka_jk_build_noalloc(h, _lib.ptr(D), _lib.ptr(J1), _lib.ptr(K1),
                                                _lib.WANT_J | _lib.TIME, sptr))
