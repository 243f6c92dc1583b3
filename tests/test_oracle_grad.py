"""Pins of the oracle's nuclear-derivative restatement (CPU).

The oracle's derivative contraction (reference _eri_rys_cpu.pyx:4316-4656)
and its ordered-pair gradient driver (mcscf/nac/_dense.py:195-250,
:1513-1839) are checked against what the mathematics fixes:

* per centre: moving one shell centre by +-h changes sum bar * (ab|cd) by the
  contracted derivative of that centre (five-point central differences of
  the oracle's own tiles), for quartets of four distinct shells up to f;
* the full functional L(R) = Tr(WJ J(D)) + Tr(WK K(D)) at fixed matrices:
  its gradient equals central differences of L over displaced geometries
  (dense oracle ERIs), and sums to zero over the atoms (translational
  invariance).
A wrong sign, a dropped -l*I(l-1) term, a swapped exponent or a missing pair
folding fails one of these.
"""

import dataclasses

import numpy as np
import pytest

from asuka_amd.basis import build_ao_basis_cart
from asuka_amd.molecule import water_monomer


def _fd5(f, h):
    return (f(-2) - 8 * f(-1) + 8 * f(1) - f(2)) / (12 * h)


def test_eri_deriv_each_centre_against_shell_displacement(oracle):
    mol = water_monomer("cc-pVTZ")
    ao, _ = build_ao_basis_cart(mol)
    ctx = oracle.make_context(ao)
    L = np.asarray(ao.shell_l)
    spA, spB = ctx.sp_A, ctx.sp_B
    rng = np.random.default_rng(5)
    # quartets of four distinct shells, covering f on some centre
    cands = [(ab, cd) for ab in range(ctx.nsp) for cd in range(ctx.nsp)
             if len({int(spA[ab]), int(spB[ab]), int(spA[cd]), int(spB[cd])}) == 4]
    f_q = [q for q in cands if 3 in (L[spA[q[0]]], L[spB[q[0]]], L[spA[q[1]]], L[spB[q[1]]])]
    picks = [cands[i] for i in rng.choice(len(cands), 4, replace=False)] + [f_q[i] for i in
                                                                           rng.choice(len(f_q), 2, replace=False)]
    h = 1e-4
    for ab, cd in picks:
        shells = [int(spA[ab]), int(spB[ab]), int(spA[cd]), int(spB[cd])]
        t0 = oracle.eri_tile(ctx, ab, cd)
        bar = rng.standard_normal(t0.shape)
        d = oracle.eri_deriv(ctx, ab, cd, bar)
        for c, sh in enumerate(shells):
            for ax in range(3):
                def f(s):
                    xyz = ao.shell_cxyz.copy()
                    xyz[sh, ax] += s * h
                    c2 = oracle.make_context(dataclasses.replace(ao, shell_cxyz=xyz))
                    return float((bar * oracle.eri_tile(c2, ab, cd)).sum())
                fd = _fd5(f, h)
                scale = max(1.0, abs(fd))
                assert abs(d[c, ax] - fd) < 1e-7 * scale, (ab, cd, c, ax, d[c, ax], fd)


@pytest.mark.slow
def test_grad2e_matches_finite_differences(oracle):
    mol = water_monomer("cc-pVDZ")
    ao, _ = build_ao_basis_cart(mol)
    ctx = oracle.make_context(ao)
    nao = ao.nao
    rng = np.random.default_rng(11)
    C = rng.standard_normal((nao, 5))
    D = C @ C.T
    WJ = rng.standard_normal((nao, nao))
    WJ = WJ + WJ.T
    WK = rng.standard_normal((nao, nao))
    WK = WK + WK.T
    g = oracle.grad2e(ctx, D, WJ, WK, ao.shell_atom, mol.natm, hp=True, threads=0)
    assert np.abs(g.sum(axis=0)).max() < 1e-9 * np.abs(g).max()

    R0 = mol.coords_bohr.copy()

    def L(R):
        a, _ = build_ao_basis_cart(mol.with_coords(R))
        eri = oracle.dense_eri(oracle.make_context(a))
        J, K = oracle.dense_JK(eri, D)
        return float((WJ * J).sum() + (WK * K).sum())

    h = 1e-4
    for atom, ax in [(0, 0), (1, 2), (2, 1)]:
        def f(s):
            R = R0.copy()
            R[atom, ax] += s * h
            return L(R)
        fd = _fd5(f, h)
        assert abs(g[atom, ax] - fd) < 1e-8 * np.abs(g).max(), (atom, ax, g[atom, ax], fd)


def test_grad2e_bilinear(oracle):
    """L = Tr(WJ J(D)) + Tr(WK K(D)) is bilinear in (W, D): halving W and
    doubling D leaves the gradient unchanged, and the J and K parts add."""
    mol = water_monomer("cc-pVDZ")
    ao, _ = build_ao_basis_cart(mol)
    ctx = oracle.make_context(ao)
    nao = ao.nao
    rng = np.random.default_rng(3)
    C = rng.standard_normal((nao, 4))
    D = C @ C.T
    g1 = oracle.grad2e(ctx, D, 0.5 * D, -0.25 * D, ao.shell_atom, mol.natm, hp=False, threads=0)
    g2 = oracle.grad2e(ctx, 2.0 * D, 0.25 * D, -0.125 * D, ao.shell_atom, mol.natm, hp=False, threads=0)
    assert np.abs(g1 - g2).max() < 1e-10 * np.abs(g1).max()
    gJ = oracle.grad2e(ctx, D, 0.5 * D, None, ao.shell_atom, mol.natm, hp=False, threads=0)
    gK = oracle.grad2e(ctx, D, None, -0.25 * D, ao.shell_atom, mol.natm, hp=False, threads=0)
    assert np.abs(gJ + gK - g1).max() < 1e-10 * np.abs(g1).max()


def test_grad2e_tasks_over_all_ordered_pairs_equals_grad2e(oracle):
    """The task-list driver (the gradient CPU baseline's) sums the same per-quartet
    contributions as the full ordered-pair driver."""
    import numpy as np

    from asuka_amd.basis import build_ao_basis_cart
    from asuka_amd.molecule import water_monomer

    mol = water_monomer("cc-pVDZ")
    ao, _ = build_ao_basis_cart(mol)
    octx = oracle.make_context(ao)
    C = np.random.default_rng(5).standard_normal((ao.nao, 5))
    D = C @ C.T
    nsp = len(oracle.shell_pairs_l_order(ao)[0])
    ab = np.repeat(np.arange(nsp), nsp)
    cd = np.tile(np.arange(nsp), nsp)
    sa = np.asarray(ao.shell_atom, dtype=np.int32)
    g1 = oracle.grad2e(octx, D, 0.5 * D, -0.25 * D, sa, mol.natm, hp=False)
    g2 = oracle.grad2e_tasks(octx, D, 0.5 * D, -0.25 * D, sa, mol.natm, ab, cd, hp=False)
    assert np.abs(g1 - g2).max() < 1e-11 * max(1.0, np.abs(g1).max())
    # a subset: half the tasks plus the other half equals the whole
    h = ab.size // 2
    g3 = oracle.grad2e_tasks(octx, D, 0.5 * D, -0.25 * D, sa, mol.natm, ab[:h], cd[:h], hp=False)
    g4 = oracle.grad2e_tasks(octx, D, 0.5 * D, -0.25 * D, sa, mol.natm, ab[h:], cd[h:], hp=False)
    assert np.abs(g3 + g4 - g2).max() < 1e-11 * max(1.0, np.abs(g1).max())
