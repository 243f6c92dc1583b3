"""GPU parity of the two-density (UHF/ROHF) build, asuka_jk_build_multi.

* against the oracle's golden J/K (no screening): Da = D, Db = -0.35 D, so
  (Ja, Ka) must equal the golden pair and (Jb, Kb) -0.35 times it, to 1e-10
  absolute; a kernel that digested any density block of the second set from
  the wrong matrix fails the scaled comparison;
* against two one-density builds with unrelated densities, at the same
  quartet set (density screening off), to 1e-12 absolute, for a cc-pVDZ
  cluster, a cc-pVTZ cluster (f shells: cooperative and generic kernels) and
  (H2O)8 cc-pVTZ (520 AOs; class pairs whose two accumulator sets exceed the
  workgroup's LDS take the global-atomic digestion);
* want flags, output buffer reuse, mixed precision and argument errors.
"""

import os

import numpy as np
import pytest

from asuka_amd.basis import build_ao_basis_cart
from asuka_amd.molecule import water_cluster

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))


def _ctx(mol, **kw):
    from asuka_amd.hf.direct_jk import make_direct_jk_context

    ao, _ = build_ao_basis_cart(mol)
    return ao, make_direct_jk_context(ao, atom_coords_bohr=mol.coords_bohr, atom_charges=mol.charges, **kw)


def _density(nao, nocc, seed):
    C = np.random.default_rng(seed).standard_normal((nao, nocc))
    C, _ = np.linalg.qr(C)
    return C @ C.T


@pytest.fixture(scope="module")
def jk_golden():
    return dict(np.load(os.path.join(HERE, "golden", "jk_golden.npz"), allow_pickle=False))


def _amax(x, ref):
    return float(np.max(np.abs(x.cpu().numpy() - ref)))


@pytest.mark.parametrize("case", ["h2o2_tz", "h2o4_dz"])
def test_multi_against_golden(gpu, jk_golden, case):
    from asuka_amd.hf.direct_jk import direct_JK_multi

    n, seed, basis = {"h2o2_tz": (2, 2, "cc-pVTZ"), "h2o4_dz": (4, 4, "cc-pVDZ")}[case]
    mol = water_cluster(n, seed, basis)
    D = jk_golden[f"{case}_D"]
    J0, K0 = jk_golden[f"{case}_J_full"], jk_golden[f"{case}_K_full"]
    _, ctx = _ctx(mol, eps_schwarz=0.0, eps_density=0.0)
    c = -0.35
    Ja, Ka, Jb, Kb = direct_JK_multi(ctx, D, c * D)
    assert _amax(Ja, J0) < 1e-10 and _amax(Ka, K0) < 1e-10
    assert _amax(Jb, c * J0) < 1e-10 and _amax(Kb, c * K0) < 1e-10


@pytest.mark.parametrize("n,seed,basis", [(4, 4, "cc-pVDZ"), (2, 3, "cc-pVTZ"), (8, 2, "cc-pVTZ")])
def test_multi_matches_two_single_builds(gpu, n, seed, basis):
    from asuka_amd.hf.direct_jk import direct_JK, direct_JK_multi

    mol = water_cluster(n, seed, basis)
    ao, ctx = _ctx(mol, eps_schwarz=1e-12, eps_density=0.0)
    Da = _density(ao.nao, 5 * n, 31)
    Db = 0.7 * _density(ao.nao, 4 * n, 32)
    Ja, Ka, Jb, Kb = direct_JK_multi(ctx, Da, Db)
    Ja1, Ka1 = direct_JK(ctx, Da)
    Jb1, Kb1 = direct_JK(ctx, Db)
    for x, y in ((Ja, Ja1), (Ka, Ka1), (Jb, Jb1), (Kb, Kb1)):
        assert float((x - y).abs().max()) < 1e-12
    assert float((Ja - Ja.T).abs().max()) == 0.0 and float((Kb - Kb.T).abs().max()) == 0.0


def test_multi_flags_mixed_and_errors(gpu):
    from asuka_amd.hf.direct_jk import direct_JK_multi

    torch = gpu
    mol = water_cluster(2, 3, "cc-pVDZ")
    ao, ctx = _ctx(mol)
    Da = _density(ao.nao, 10, 41)
    Db = _density(ao.nao, 9, 42)
    Ja, Ka, Jb, Kb = direct_JK_multi(ctx, Da, Db)
    Ja1, Ka1, Jb1, Kb1 = direct_JK_multi(ctx, Da, Db, want_K=False)
    assert Ka1 is None and Kb1 is None
    assert float((Ja1 - Ja).abs().max()) < 1e-13 and float((Jb1 - Jb).abs().max()) < 1e-13
    Ja2, Ka2, Jb2, Kb2 = direct_JK_multi(ctx, Da, Db, want_J=False)
    assert Ja2 is None and Jb2 is None
    assert float((Ka2 - Ka).abs().max()) < 1e-13 and float((Kb2 - Kb).abs().max()) < 1e-13
    assert direct_JK_multi(ctx, Da, Db, want_J=False, want_K=False) == (None, None, None, None)
    # caller buffers are overwritten, not accumulated into
    bufs = tuple(torch.full((ao.nao, ao.nao), 7.0, dtype=torch.float64, device="cuda") for _ in range(4))
    out = direct_JK_multi(ctx, Da, Db, out=bufs)
    for x, y in zip(out, (Ja, Ka, Jb, Kb)):
        assert float((x - y).abs().max()) < 1e-13
    # mixed precision: within the bound of the one-density mixed build
    m = direct_JK_multi(ctx, Da, Db, mixed_precision=True)
    for x, y in zip(m, (Ja, Ka, Jb, Kb)):
        assert float((x - y).abs().max()) < 1e-8
    with pytest.raises(ValueError, match="must be"):
        direct_JK_multi(ctx, Da, np.zeros((3, 3)))
    with pytest.raises(ValueError, match="distinct"):
        direct_JK_multi(ctx, Da, Db, out=(bufs[0], bufs[1], bufs[0], bufs[3]))
