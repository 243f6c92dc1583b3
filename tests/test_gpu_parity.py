"""GPU parity: the HIP path through the C ABI against the CPU oracle.

Bars (FP64 mode): ERI tiles, Schwarz bounds and 1e integrals to 1e-12
absolute; J and K to 1e-10 absolute; SCF energy to 1e-9 Eh.
"""

import os

import numpy as np
import pytest

from asuka_amd.basis import build_ao_basis_cart
from asuka_amd.molecule import water_cluster, water_monomer

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))


def _ctx(mol, **kw):
    from asuka_amd.hf.direct_jk import make_direct_jk_context

    ao, _ = build_ao_basis_cart(mol)
    return ao, make_direct_jk_context(ao, atom_coords_bohr=mol.coords_bohr, atom_charges=mol.charges, **kw)


def _density(nao, nocc=10, seed=11):
    C = np.random.default_rng(seed).standard_normal((nao, nocc))
    C, _ = np.linalg.qr(C)
    return C @ C.T


@pytest.fixture(scope="module")
def jk_golden():
    return dict(np.load(os.path.join(HERE, "golden", "jk_golden.npz"), allow_pickle=False))


def test_rys_roots_device_against_golden(gpu, golden):
    import ctypes

    from asuka_amd import _lib

    torch = gpu
    T = torch.as_tensor(golden["T"], device="cuda")
    for n in range(1, 8):
        x = torch.empty((T.numel(), n), dtype=torch.float64, device="cuda")
        w = torch.empty_like(x)
        _lib.check(_lib.lib().asuka_rys_roots_device(n, T.numel(), _lib.ptr(T), _lib.ptr(x), _lib.ptr(w),
                                                      ctypes.c_void_p(_lib.stream_handle())))
        np.testing.assert_allclose(x.cpu().numpy(), golden["x"][n, :, :n], rtol=5e-15, atol=0)
        np.testing.assert_allclose(w.cpu().numpy(), golden["w"][n, :, :n], rtol=5e-15, atol=0)


@pytest.mark.parametrize("basis", ["cc-pVDZ", "cc-pVTZ"])
def test_eri_tiles_all_classes(gpu, oracle, basis):
    """Every shell quartet class of H2O (s..f): tiles vs the oracle."""
    from asuka_amd.hf.direct_jk import eri_tiles

    mol = water_monomer(basis)
    ao, ctx = _ctx(mol, eps_schwarz=0.0)
    octx = oracle.make_context(ao)
    assert np.array_equal(octx.sp_A, ctx.sp_A) and np.array_equal(octx.sp_B, ctx.sp_B)
    L = ao.shell_l
    pcls = np.array([(int(L[a]), int(L[b])) for a, b in zip(ctx.sp_A, ctx.sp_B)])
    # one (or, for DZ, every) quartet per ordered class pair, both orientations
    rng = np.random.default_rng(5)
    ab_l, cd_l = [], []
    for i in range(ctx.nsp):
        for j in range(ctx.nsp):
            ab_l.append(i)
            cd_l.append(j)
    ab = np.array(ab_l)
    cd = np.array(cd_l)
    if basis == "cc-pVTZ":
        keys = [tuple(pcls[a]) + tuple(pcls[c]) for a, c in zip(ab, cd)]
        first = {}
        for k_, (a, c) in zip(keys, zip(ab, cd)):
            first.setdefault(k_, []).append((a, c))
        pick = []
        for k_, lst in first.items():
            for idx in rng.choice(len(lst), size=min(3, len(lst)), replace=False):
                pick.append(lst[idx])
        ab = np.array([p[0] for p in pick])
        cd = np.array([p[1] for p in pick])
    tiles = eri_tiles(ctx, ab, cd)
    worst = 0.0
    for t, a, c in zip(tiles, ab, cd):
        ref = oracle.eri_tile(octx, int(a), int(c), hp=True)
        worst = max(worst, float(np.max(np.abs(t - ref))))
    assert worst < 1e-12, worst


def test_schwarz_and_int1e(gpu, oracle):
    from asuka_amd.frontend import build_int1e

    for basis in ("cc-pVDZ", "cc-pVTZ"):
        mol = water_cluster(2, 2, basis)
        ao, ctx = _ctx(mol)
        octx = oracle.make_context(ao)
        Q = oracle.schwarz(octx, hp=True)
        assert np.max(np.abs(ctx.Q_sp - Q)) < 1e-12
        ints = build_int1e(ctx)
        S, T, V = oracle.int1e(ao, mol.coords_bohr, mol.charges)
        assert np.max(np.abs(ints.S.cpu().numpy() - S)) < 1e-12
        assert np.max(np.abs(ints.T.cpu().numpy() - T)) < 1e-12
        assert np.max(np.abs(ints.V.cpu().numpy() - V)) < 1e-11


def test_jk_h2o_dz_against_dense(gpu, oracle):
    """Config: H2O cc-pVDZ, J/K from the dense FP64 ERIs (hf/dense_jk.py:55-81)."""
    from asuka_amd.hf.direct_jk import direct_JK

    mol = water_monomer("cc-pVDZ")
    ao, ctx = _ctx(mol, eps_schwarz=0.0, eps_density=0.0)
    octx = oracle.make_context(ao)
    eri = oracle.dense_eri(octx, hp=True)
    rng = np.random.default_rng(42)
    D = rng.normal(size=(ao.nao, ao.nao))
    D = D + D.T
    J0, K0 = oracle.dense_JK(eri, D)
    J, K = direct_JK(ctx, D)
    assert np.max(np.abs(J.cpu().numpy() - J0)) < 1e-10
    assert np.max(np.abs(K.cpu().numpy() - K0)) < 1e-10


@pytest.mark.parametrize("case", ["h2o2_tz", "h2o4_dz"])
def test_jk_clusters_against_oracle(gpu, jk_golden, case):
    from asuka_amd.hf.direct_jk import direct_JK

    n, seed, basis = {"h2o2_tz": (2, 2, "cc-pVTZ"), "h2o4_dz": (4, 4, "cc-pVDZ")}[case]
    mol = water_cluster(n, seed, basis)
    D = jk_golden[f"{case}_D"]
    ao, ctx = _ctx(mol, eps_schwarz=1e-12)
    assert np.max(np.abs(ctx.Q_sp - jk_golden[f"{case}_Q"])) < 1e-12
    J, K = direct_JK(ctx, D)
    assert np.max(np.abs(J.cpu().numpy() - jk_golden[f"{case}_J_scr"])) < 1e-10
    assert np.max(np.abs(K.cpu().numpy() - jk_golden[f"{case}_K_scr"])) < 1e-10
    ao, ctx0 = _ctx(mol, eps_schwarz=0.0, eps_density=0.0)
    J, K = direct_JK(ctx0, D)
    assert np.max(np.abs(J.cpu().numpy() - jk_golden[f"{case}_J_full"])) < 1e-10
    assert np.max(np.abs(K.cpu().numpy() - jk_golden[f"{case}_K_full"])) < 1e-10


def test_fock_matches_h_plus_j_minus_half_k(gpu):
    from asuka_amd.hf.direct_jk import direct_fock_rhf, direct_JK

    torch = gpu
    for basis in ("cc-pVDZ", "cc-pVTZ"):
        mol = water_cluster(2, 3, basis)
        ao, ctx = _ctx(mol)
        D = _density(ao.nao)
        rng = np.random.default_rng(1)
        h = rng.normal(size=D.shape)
        h = h + h.T
        J, K = direct_JK(ctx, D)
        F = direct_fock_rhf(ctx, D, h)
        ref = torch.as_tensor(h, device="cuda") + J - 0.5 * K
        assert float((F - ref).abs().max()) < 1e-11


def test_scf_energy_h2o_dz_against_oracle(gpu, oracle):
    from asuka_amd.frontend import run_hf_df

    mol = water_monomer("cc-pVDZ")
    res = run_hf_df(mol, two_e_backend="direct", eps_schwarz=1e-12)
    assert res.scf.converged
    ao, _ = build_ao_basis_cart(mol)
    octx = oracle.make_context(ao)
    eri = oracle.dense_eri(octx, hp=True)
    S, T, V = oracle.int1e(ao, mol.coords_bohr, mol.charges)
    h = T + V

    def fock(D):
        J, K = oracle.dense_JK(eri, D)
        return 0.5 * ((h + J - 0.5 * K) + (h + J - 0.5 * K).T)

    ref = oracle.rhf(S, h, fock, mol.nelectron, mol.energy_nuc())
    assert abs(res.scf.e_tot - ref["e_tot"]) < 1e-9


def test_want_flags_and_errors(gpu):
    from asuka_amd.hf.direct_jk import direct_JK

    mol = water_monomer("cc-pVDZ")
    ao, ctx = _ctx(mol)
    D = _density(ao.nao, nocc=5)
    J, K = direct_JK(ctx, D)
    J1, K1 = direct_JK(ctx, D, want_K=False)
    assert K1 is None and float((J1 - J).abs().max()) < 1e-13
    J2, K2 = direct_JK(ctx, D, want_J=False)
    assert J2 is None and float((K2 - K).abs().max()) < 1e-13
    assert direct_JK(ctx, D, want_J=False, want_K=False) == (None, None)
    with pytest.raises(ValueError, match="must be"):
        direct_JK(ctx, np.zeros((ao.nao, ao.nao + 1)))
    with pytest.raises(ValueError, match="must be"):
        direct_JK(ctx, np.zeros((3, 3)))


def test_unsupported_angular_momentum(gpu):
    from asuka_amd.basis import BasisCartSoA
    from asuka_amd.hf.direct_jk import make_direct_jk_context

    b = BasisCartSoA(np.zeros((1, 3)), np.zeros(1, np.int32), np.ones(1, np.int32), np.array([4], np.int32),
                     np.zeros(1, np.int32), np.ones(1), np.ones(1))
    with pytest.raises(NotImplementedError):
        make_direct_jk_context(b)


def test_config2_size_properties(gpu):
    """(H2O)8 cc-pVTZ (520 AOs): size-independent properties at full size:
    symmetry, linearity in D (screening off), Tr(D J(D)) > 0, K positive
    semidefinite for D = C C^T, and J/K from Fock consistency."""
    from asuka_amd.hf.direct_jk import direct_JK

    torch = gpu
    mol = water_cluster(8, 2, "cc-pVTZ")
    ao, ctx = _ctx(mol, eps_schwarz=1e-12, eps_density=0.0)
    assert ao.nao == 520
    D1 = _density(ao.nao, 40, seed=21)
    D2 = _density(ao.nao, 40, seed=22)
    J1, K1 = direct_JK(ctx, D1)
    J2, K2 = direct_JK(ctx, D2)
    J12, K12 = direct_JK(ctx, D1 + 0.5 * D2)
    assert float((J12 - J1 - 0.5 * J2).abs().max()) < 1e-10
    assert float((K12 - K1 - 0.5 * K2).abs().max()) < 1e-10
    assert float((J1 - J1.T).abs().max()) == 0.0
    D1t = torch.as_tensor(D1, device="cuda")
    assert float((D1t * J1).sum()) > 0
    ev = torch.linalg.eigvalsh(K1)
    assert float(ev.min()) > -1e-9 * float(ev.max())


@pytest.mark.parametrize("case", ["h2o4_dz", "h2o2_tz"])
def test_mixed_precision_jk(gpu, jk_golden, case):
    """Mixed FP32/FP64 mode. With the default bound cut (1e-5) J and K stay
    within 1e-8 of the FP64 oracle; with every quartet forced onto the FP32
    kernels (cut 1e3) they stay within 1e-5, which no index or layout mistake
    in the FP32 instantiations would meet, and differ from FP64 (so the FP32
    path did run). The cc-pVTZ case runs the FP32 instantiations of the f-shell
    plans."""
    from asuka_amd.hf.direct_jk import direct_JK

    n, seed, basis = {"h2o4_dz": (4, 4, "cc-pVDZ"), "h2o2_tz": (2, 2, "cc-pVTZ")}[case]
    mol = water_cluster(n, seed, basis)
    D = jk_golden[f"{case}_D"]
    Jr, Kr = jk_golden[f"{case}_J_scr"], jk_golden[f"{case}_K_scr"]
    ao, ctx = _ctx(mol, eps_schwarz=1e-12)
    J, K = direct_JK(ctx, D, mixed_precision=True)
    assert np.max(np.abs(J.cpu().numpy() - Jr)) < 1e-8
    assert np.max(np.abs(K.cpu().numpy() - Kr)) < 1e-8
    ao, ctx32 = _ctx(mol, eps_schwarz=1e-12, tau_mixed=1e3)
    J, K = direct_JK(ctx32, D, mixed_precision=True)
    dJ = np.max(np.abs(J.cpu().numpy() - Jr))
    dK = np.max(np.abs(K.cpu().numpy() - Kr))
    assert dJ < 1e-5 and dK < 1e-5
    assert dJ > 1e-13 and dK > 1e-13
    J64, K64 = direct_JK(ctx32, D)
    assert np.max(np.abs(J64.cpu().numpy() - Jr)) < 1e-10


def test_mixed_precision_scf_energy(gpu):
    """RHF energy of (H2O)2 cc-pVDZ with the mixed-precision build: with the
    driver's switch to FP64 builds once ||D_new - D||_F < 1e-3 (graph and
    eager drivers) the converged energy is the FP64 one to 1e-9 Eh; with the
    builds mixed to the end it stays within 1e-6 Eh."""
    from asuka_amd.frontend import run_hf_df

    mol = water_cluster(2, 3, "cc-pVDZ")
    e64 = run_hf_df(mol, two_e_backend="direct", eps_schwarz=1e-12, mixed_precision=False)
    assert e64.scf.converged
    for use_graph in (True, False):
        prof = {}
        emx = run_hf_df(mol, two_e_backend="direct", eps_schwarz=1e-12, mixed_precision=True, use_graph=use_graph,
                        profile=prof)
        assert emx.scf.converged and prof["scf"]["graph"] is use_graph
        assert prof["scf"]["fp64_from_cycle"] is not None and prof["scf"]["fp64_from_cycle"] > 1
        assert abs(emx.scf.e_tot - e64.scf.e_tot) < 1e-9, (use_graph, emx.scf.e_tot - e64.scf.e_tot)
    import os

    os.environ["ASUKA_SCF_MIXED_UNTIL_DM"] = "0"
    try:
        prof = {}
        emx = run_hf_df(mol, two_e_backend="direct", eps_schwarz=1e-12, mixed_precision=True, profile=prof)
    finally:
        del os.environ["ASUKA_SCF_MIXED_UNTIL_DM"]
    assert emx.scf.converged and prof["scf"]["fp64_from_cycle"] is None
    assert abs(emx.scf.e_tot - e64.scf.e_tot) < 1e-6


def test_stats_match_efm_model(gpu):
    """asuka_jk_stats after a counted build equals the EFM-v1 model evaluated
    in Python on the per-class counters (asuka_amd/metrics.py)."""
    import ctypes

    from asuka_amd import _lib, metrics
    from asuka_amd.hf.direct_jk import direct_JK

    torch = gpu
    mol = water_cluster(2, 3, "cc-pVTZ")
    ao, ctx = _ctx(mol)
    D = torch.tensor(_density(ao.nao, 10, 5), device="cuda")
    direct_JK(ctx, D)  # warm
    J = torch.empty_like(D)
    K = torch.empty_like(D)
    L = _lib.lib()
    sptr = ctypes.c_void_p(_lib.stream_handle())
    _lib.check(L.asuka_jk_build(ctypes.c_void_p(ctx.handle), _lib.ptr(D), _lib.ptr(J), _lib.ptr(K),
                                _lib.WANT_J | _lib.WANT_K | _lib.MIXED | _lib.COUNT | _lib.TIME, sptr))
    st = ctx.stats()
    ncp = len(ctx.class_stats())
    rows = np.zeros((ncp, 16), dtype=np.int64)
    L.asuka_jk_last_counts(ctypes.c_void_p(ctx.handle), rows.ctypes.data_as(_lib._i64p), ncp, sptr)
    cls = [s["class"] for s in ctx.class_stats()]
    work = metrics.build_work([(c, rows[i, :8].tolist(), int(rows[i, 8])) for i, c in enumerate(cls)])
    assert st["class_pairs"] == ncp and st["schwarz_quartets"] == ctx.ntasks
    assert st["quartets"] == work.quartets and st["prim_quartets"] == work.prim_quartets
    assert abs(st["flops_efm"] - work.flops) <= 1e-9 * work.flops
    f32 = sum(metrics.flop_eri(*c, int(rows[i, 10])) for i, c in enumerate(cls))
    assert abs(st["flops_efm_fp32"] - f32) <= 1e-9 * max(1.0, f32)
    assert 0 < st["quartets_fp32"] < st["quartets"] <= st["schwarz_quartets"]
    assert st["last_ms"] > 0.0
