"""Deterministic builds (ASUKA_JK_DETERMINISTIC): every J/K/F contribution is
rounded to a multiple of 2^-44 and summed as a 64-bit integer, so the result
does not depend on the order of the atomic adds.

* J and K of the golden cases (register, cooperative and generic kernels)
  equal the oracle's to 1e-10, and three builds on the default 4 streams are
  bitwise equal;
* (H2O)8 cc-pVTZ (520 AOs: LDS row accumulators, ket-position J buffer,
  permuted K flush, global-atomic digestion of the class pairs whose
  accumulators exceed the LDS): bitwise equal over three builds, FP64 and
  mixed, and within 1e-9 of the ordinary build;
* the Fock build and the two-density build (density screening off): bitwise
  equal over three builds, equal to h + J - K/2 and to the one-density builds
  to 1e-10;
* the batch builds reject the flag.
"""

import os

import numpy as np
import pytest

from asuka_amd.basis import build_ao_basis_cart
from asuka_amd.molecule import water_cluster

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
CASES = {"h2o4_dz": (4, 4, "cc-pVDZ"), "h2o2_tz": (2, 2, "cc-pVTZ")}


@pytest.fixture(scope="module")
def jk_golden():
    return dict(np.load(os.path.join(HERE, "golden", "jk_golden.npz"), allow_pickle=False))


def _ctx(mol, **kw):
    from asuka_amd.hf.direct_jk import make_direct_jk_context

    ao, _ = build_ao_basis_cart(mol)
    return ao, make_direct_jk_context(ao, atom_coords_bohr=mol.coords_bohr, atom_charges=mol.charges, **kw)


def _density(nao, nocc, seed):
    C = np.random.default_rng(seed).standard_normal((nao, nocc))
    C, _ = np.linalg.qr(C)
    return C @ C.T


def _bitwise_equal(torch, runs):
    return all(torch.equal(r, runs[0]) for r in runs[1:])


@pytest.mark.parametrize("case", sorted(CASES))
def test_deterministic_against_golden_and_repeatable(gpu, jk_golden, case):
    from asuka_amd.hf.direct_jk import direct_JK

    torch = gpu
    n, seed, basis = CASES[case]
    print(f"\n[det] {case}", flush=True)
    ao, ctx = _ctx(water_cluster(n, seed, basis), eps_schwarz=1e-12)
    D = jk_golden[f"{case}_D"]
    Jr, Kr = jk_golden[f"{case}_J_scr"], jk_golden[f"{case}_K_scr"]
    runs = [direct_JK(ctx, D, mixed_precision=False, deterministic=True) for _ in range(3)]
    J, K = runs[0]
    assert np.max(np.abs(J.cpu().numpy() - Jr)) < 1e-10
    assert np.max(np.abs(K.cpu().numpy() - Kr)) < 1e-10
    assert _bitwise_equal(torch, [r[0] for r in runs]) and _bitwise_equal(torch, [r[1] for r in runs])


def test_deterministic_h2o8_tz_fp64_and_mixed(gpu):
    from asuka_amd.hf.direct_jk import direct_JK

    torch = gpu
    print("\n[det] (H2O)8 cc-pVTZ", flush=True)
    ao, ctx = _ctx(water_cluster(8, 2, "cc-pVTZ"), eps_schwarz=1e-12)
    D = _density(ao.nao, 40, 1002)
    J0, K0 = direct_JK(ctx, D, mixed_precision=False)
    for mixed in (False, True):
        runs = [direct_JK(ctx, D, mixed_precision=mixed, deterministic=True) for _ in range(3)]
        assert _bitwise_equal(torch, [r[0] for r in runs]), mixed
        assert _bitwise_equal(torch, [r[1] for r in runs]), mixed
        tol = 1e-9 if not mixed else 1e-8
        assert float((runs[0][0] - J0).abs().max()) < tol and float((runs[0][1] - K0).abs().max()) < tol, mixed
        assert float((runs[0][0] - runs[0][0].T).abs().max()) == 0.0


def test_deterministic_fock_and_multi(gpu):
    from asuka_amd.frontend import build_int1e
    from asuka_amd.hf.direct_jk import direct_fock_rhf, direct_JK, direct_JK_multi

    torch = gpu
    print("\n[det] (H2O)4 cc-pVDZ Fock and two-density builds", flush=True)
    # density screening off: the two-density build screens by max(Da, Db), so
    # only then is its quartet set that of the one-density builds
    ao, ctx = _ctx(water_cluster(4, 4, "cc-pVDZ"), eps_schwarz=1e-12, eps_density=0.0)
    h = build_int1e(ctx).hcore
    Da = _density(ao.nao, 20, 31)
    Db = _density(ao.nao, 18, 32)
    F = [direct_fock_rhf(ctx, Da, h, mixed_precision=False, deterministic=True) for _ in range(3)]
    assert _bitwise_equal(torch, F)
    J, K = direct_JK(ctx, Da, mixed_precision=False)
    hd = torch.as_tensor(np.asarray(h) if not isinstance(h, torch.Tensor) else h, device="cuda")
    assert float((F[0] - (hd + J - 0.5 * K)).abs().max()) < 1e-10
    M = [direct_JK_multi(ctx, Da, Db, mixed_precision=False, deterministic=True) for _ in range(3)]
    for i in range(4):
        assert _bitwise_equal(torch, [m[i] for m in M]), i
    Jb, Kb = direct_JK(ctx, Db, mixed_precision=False)
    for got, want in zip(M[0], (J, K, Jb, Kb)):
        assert float((got - want).abs().max()) < 1e-10


def test_batch_builds_reject_deterministic(gpu):
    import ctypes

    from asuka_amd import _lib
    from asuka_amd.hf.direct_jk import make_direct_jk_batch_context

    torch = gpu
    aos = [build_ao_basis_cart(water_cluster(n, 7 + n, "cc-pVDZ"))[0] for n in (1, 2)]
    bc = make_direct_jk_batch_context(aos, eps_schwarz=1e-12)
    Ds = [torch.as_tensor(_density(a.nao, 5, 3), device="cuda") for a in aos]
    Js = [torch.empty_like(d) for d in Ds]
    arr = lambda xs: (ctypes.c_void_p * len(xs))(*[x.data_ptr() for x in xs])  # noqa: E731
    flags = _lib.WANT_J | _lib.DETERMINISTIC
    rc = _lib.lib().asuka_jk_batch_build(ctypes.c_void_p(bc.handle), arr(Ds), arr(Js), None, flags,
                                         ctypes.c_void_p(_lib.stream_handle()))
    assert rc != 0
    assert "DETERMINISTIC" in _lib.lib().asuka_jk_last_error().decode()
