"""Pins of the CPU oracle (no GPU). Each test checks the oracle against something
it does not share code with: 40-digit golden values, an independent textbook
McMurchie-Davidson implementation, closed forms or exact invariants.
"""

import math

import numpy as np
import pytest

import mmd_ref
from asuka_amd.basis import BasisCartSoA, build_ao_basis_cart, cartesian_components, ncart
from asuka_amd.molecule import water_cluster, water_monomer


def _raw_basis(shells):
    """BasisCartSoA from (l, center, [(exp, coef), ...]) with coefficients used as given."""
    cxyz, ps, npr, ls, ao, pe, pc = [], [], [], [], [], [], []
    a = 0
    for l, c, prims in shells:
        cxyz.append(c)
        ps.append(len(pe))
        npr.append(len(prims))
        ls.append(l)
        ao.append(a)
        a += ncart(l)
        for e, k in prims:
            pe.append(e)
            pc.append(k)
    return BasisCartSoA(np.asarray(cxyz, float).reshape(-1, 3), np.asarray(ps, np.int32), np.asarray(npr, np.int32),
                        np.asarray(ls, np.int32), np.asarray(ao, np.int32), np.asarray(pe, float),
                        np.asarray(pc, float))


# ---------------------------------------------------------------- Boys / Rys


def test_boys_against_golden(oracle, golden):
    for i, T in enumerate(golden["T"]):
        F = oracle.boys_list(T, 16)
        np.testing.assert_allclose(F, golden["boys"][i], rtol=2e-11, atol=0)


def test_boys_special_values(oracle):
    F = oracle.boys_list(0.0, 10)
    np.testing.assert_allclose(F, [1.0 / (2 * m + 1) for m in range(11)], rtol=1e-15)
    for T in (0.3, 2.0, 9.0, 40.0):
        f0 = 0.5 * math.sqrt(math.pi / T) * math.erf(math.sqrt(T))
        assert abs(oracle.boys_list(T, 0)[0] - f0) < 1e-15


@pytest.mark.parametrize("n", range(1, 8))
def test_rys_hp_against_golden(oracle, golden, n):
    for i, T in enumerate(golden["T"]):
        x, w = oracle.rys_roots_weights(n, T, hp=True)
        np.testing.assert_allclose(x, golden["x"][n, i, :n], rtol=1e-14)
        np.testing.assert_allclose(w, golden["w"][n, i, :n], rtol=1e-14)


def test_rys_double_mode_is_the_reference_error(oracle, golden):
    """The reference's double-precision moment algorithm degrades with n: it is
    accurate for n <= 3 and measurably off at n = 7 (the reason hp mode exists)."""
    worst = {}
    for n in (3, 7):
        e = 0.0
        for i, T in enumerate(golden["T"]):
            x, w = oracle.rys_roots_weights(n, T, hp=False)
            e = max(e, float(np.max(np.abs(x - golden["x"][n, i, :n]) / golden["x"][n, i, :n])))
        worst[n] = e
    assert worst[3] < 1e-12
    assert 1e-11 < worst[7] < 1e-7


@pytest.mark.parametrize("n", [1, 4, 7])
def test_rys_moments_reproduce_boys(oracle, golden, n):
    for i, T in enumerate(golden["T"]):
        x, w = oracle.rys_roots_weights(n, T, hp=True)
        for k in range(2 * n):
            assert abs(np.sum(w * x ** k) - golden["boys"][i][k]) <= 1e-14 * golden["boys"][i][0]


# ---------------------------------------------------------------- ERI tiles


def test_ssss_closed_form(oracle, golden):
    rng = np.random.default_rng(7)
    for _ in range(5):
        exps = rng.uniform(0.1, 5.0, 4)
        cents = rng.normal(size=(4, 3))
        b = _raw_basis([(0, cents[i], [(exps[i], 1.0)]) for i in range(4)])
        ctx = oracle.make_context(b)
        # pair index of (i, j), i >= j: i(i+1)/2 + j
        ab, cd = 1 * 2 // 2 + 0, 3 * 4 // 2 + 2
        v = oracle.eri_tile(ctx, ab, cd)[0, 0]
        a, bb, c, d = exps[1], exps[0], exps[3], exps[2]
        A, B, C, D = cents[1], cents[0], cents[3], cents[2]
        p, q = a + bb, c + d
        P, Q = (a * A + bb * B) / p, (c * C + d * D) / q
        Kab = math.exp(-a * bb / p * np.sum((A - B) ** 2))
        Kcd = math.exp(-c * d / q * np.sum((C - D) ** 2))
        T = p * q / (p + q) * np.sum((P - Q) ** 2)
        ref = 2 * math.pi ** 2.5 / (p * q * math.sqrt(p + q)) * Kab * Kcd * mmd_ref.boys_array(0, T)[0]
        assert abs(v - ref) < 1e-14 * abs(ref)


@pytest.mark.parametrize("ls", [(0, 0, 0, 0), (1, 0, 1, 0), (1, 1, 1, 1), (2, 1, 1, 0), (2, 2, 1, 1), (2, 2, 2, 2),
                                (3, 0, 2, 1), (3, 2, 3, 1), (3, 3, 2, 2), (3, 3, 3, 3)])
def test_eri_tile_against_mmd(oracle, ls):
    """Contracted (here: single-primitive) tiles vs an independent MMD code."""
    rng = np.random.default_rng(sum(l * 10 ** i for i, l in enumerate(ls)))
    exps = rng.uniform(0.2, 3.0, 4)
    cents = rng.normal(scale=0.8, size=(4, 3))
    la, lb, lc, ld = ls
    # shell order so that pairs (1,0) and (3,2) keep l-ordered orientation A=1,B=0 / C=3,D=2
    shells = [(lb, cents[1], [(exps[1], 1.0)]), (la, cents[0], [(exps[0], 1.0)]),
              (ld, cents[3], [(exps[3], 1.0)]), (lc, cents[2], [(exps[2], 1.0)])]
    b = _raw_basis(shells)
    ctx = oracle.make_context(b)
    ab = 1 * 2 // 2 + 0
    cd = 3 * 4 // 2 + 2
    assert ctx.sp_A[ab] == 1 and ctx.sp_A[cd] == 3
    tile = oracle.eri_tile(ctx, ab, cd)
    ref = mmd_ref.eri_prim_tile(la, lb, lc, ld, exps[0], exps[1], exps[2], exps[3], cents[0], cents[1], cents[2],
                                cents[3])
    scale = np.max(np.abs(ref))
    np.testing.assert_allclose(tile, ref, atol=2e-13 * max(scale, 1.0), rtol=0)


@pytest.fixture(scope="module")
def h2o_dz(oracle):
    mol = water_monomer("cc-pVDZ")
    ao, _ = build_ao_basis_cart(mol)
    ctx = oracle.make_context(ao)
    eri = oracle.dense_eri(ctx)
    return mol, ao, ctx, eri


def test_dense_eri_symmetry_and_schwarz(h2o_dz, oracle):
    mol, ao, ctx, eri = h2o_dz
    n = ao.nao
    e4 = eri.reshape(n, n, n, n)
    assert np.max(np.abs(e4 - e4.transpose(1, 0, 2, 3))) == 0.0
    assert np.max(np.abs(e4 - e4.transpose(2, 3, 0, 1))) == 0.0
    diag = np.einsum("ijij->ij", e4)
    assert np.all(diag >= 0)
    lhs = eri ** 2
    rhs = np.outer(diag.ravel(), diag.ravel())
    assert np.all(lhs <= rhs * (1 + 1e-12) + 1e-30)
    w = np.linalg.eigvalsh(0.5 * (eri + eri.T))
    assert w.min() > -1e-10 * w.max()
    Q = oracle.schwarz(ctx)
    for sp in range(ctx.nsp):
        A, B = ctx.sp_A[sp], ctx.sp_B[sp]
        a0, b0 = ao.shell_ao_start[A], ao.shell_ao_start[B]
        na, nb = ncart(ao.shell_l[A]), ncart(ao.shell_l[B])
        blk = diag[a0:a0 + na, b0:b0 + nb]
        assert abs(Q[sp] - math.sqrt(blk.max())) <= 1e-14 * max(1.0, Q[sp])


def test_direct_jk_matches_dense_einsum(h2o_dz, oracle):
    """Digestion semantics of the ordered contraction == the einsum definitions."""
    mol, ao, ctx, eri = h2o_dz
    rng = np.random.default_rng(42)
    D = rng.normal(size=(ao.nao, ao.nao))
    D = D + D.T
    J0, K0 = oracle.dense_JK(eri, D)
    J1, K1 = oracle.direct_jk(ctx, D, eps_schwarz=0.0, eps_density=0.0)
    assert np.max(np.abs(J1 - J0)) < 1e-12
    assert np.max(np.abs(K1 - K0)) < 1e-12


def test_screened_tasks_match_bruteforce(oracle):
    rng = np.random.default_rng(3)
    Q = np.exp(rng.uniform(-30, 1, 200))
    Q[rng.integers(0, 200, 10)] = 0.0
    for eps in (0.0, 1e-12, 1e-6):
        ab, cd = oracle.tasks_screened_sorted_q(Q, eps)
        got = set(zip(ab.tolist(), cd.tolist()))
        ref = {(i, j) for i in range(200) for j in range(i + 1) if Q[i] * Q[j] >= eps and Q[i] > 0 and Q[j] > 0}
        assert got == ref
        assert np.all(cd <= ab)


# ---------------------------------------------------------------- basis / 1e / SCF


def test_basis_normalisation_invariants(oracle):
    """Self-overlaps: 1 for s/p; 4 pi (2lx-1)!!(2ly-1)!!(2lz-1)!!/(2l+1)!! for l >= 2."""
    df = lambda k: 1 if k <= 0 else k * df(k - 2)
    for name in ("cc-pVDZ", "cc-pVTZ"):
        mol = water_monomer(name)
        ao, _ = build_ao_basis_cart(mol)
        S, T, V = oracle.int1e(ao, mol.coords_bohr, mol.charges)
        for sh in range(ao.nshell):
            l = int(ao.shell_l[sh])
            a0 = int(ao.shell_ao_start[sh])
            for i, (x, y, z) in enumerate(cartesian_components(l)):
                ref = 1.0 if l <= 1 else 4 * math.pi * df(2 * x - 1) * df(2 * y - 1) * df(2 * z - 1) / df(2 * l + 1)
                assert abs(S[a0 + i, a0 + i] - ref) < 1e-13


def test_int1e_against_mmd(oracle):
    rng = np.random.default_rng(11)
    cents = rng.normal(size=(3, 3))
    shells = [(3, cents[0], [(1.3, 1.0)]), (2, cents[1], [(0.7, 1.0)]), (1, cents[2], [(2.1, 1.0)]),
              (0, cents[0], [(0.4, 1.0)])]
    b = _raw_basis(shells)
    atoms = rng.normal(size=(2, 3))
    Z = np.array([8.0, 1.0])
    S, T, V = oracle.int1e(b, atoms, Z)
    offs = [0, 10, 16, 19]
    for i, (li, ci, pi) in enumerate(shells):
        for j, (lj, cj, pj) in enumerate(shells):
            s_ref = mmd_ref.overlap_prim(li, lj, pi[0][0], pj[0][0], ci, cj)
            v_ref = sum(mmd_ref.nuclear_prim(li, lj, pi[0][0], pj[0][0], ci, cj, atoms[k], Z[k]) for k in range(2))
            blk = (slice(offs[i], offs[i] + ncart(li)), slice(offs[j], offs[j] + ncart(lj)))
            np.testing.assert_allclose(S[blk], s_ref, atol=1e-13)
            np.testing.assert_allclose(V[blk], v_ref, atol=1e-12)
    # kinetic energy: T = -1/2 <a|lap|b>; s-s closed form
    a, bexp = 0.4, 1.3
    d2 = np.sum((cents[0] - cents[0]) ** 2)
    assert np.allclose(T, T.T, atol=1e-14)
    w = np.linalg.eigvalsh(T)
    assert w.min() > 0  # kinetic energy matrix is positive definite


def test_h2o_dz_rhf_energy_variational(h2o_dz, oracle):
    mol, ao, ctx, eri = h2o_dz
    S, T, V = oracle.int1e(ao, mol.coords_bohr, mol.charges)
    h = T + V

    def fock(D):
        J, K = oracle.dense_JK(eri, D)
        return 0.5 * ((h + J - 0.5 * K) + (h + J - 0.5 * K).T)

    r = oracle.rhf(S, h, fock, mol.nelectron, mol.energy_nuc())
    assert r["converged"]
    # stationarity: FDS - SDF = 0 at convergence, and E agrees with the
    # variational expression from the converged density
    D = r["dm"]
    F = fock(D)
    assert np.max(np.abs(F @ D @ S - S @ D @ F)) < 1e-6
    e_var = 0.5 * np.sum(D * (h + F)) + mol.energy_nuc()
    assert abs(e_var - r["e_tot"]) < 1e-9
    assert -76.1 < r["e_tot"] < -75.9
