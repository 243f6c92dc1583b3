"""Pins the oracle's density-fitting integrals and J/K (CPU, no GPU).

* three-centre (ab|P) and two-centre (P|Q) integrals of single primitives,
  every l up to f on three centres, against the independent McMurchie-Davidson
  ERI (tests/mmd_ref.py) with the fourth exponent 0 (the constant function);
* DF Coulomb energy of the water monomer (cc-pVDZ, even-tempered auxiliary
  set) below the exact one (the Coulomb-metric fit is variational) and within
  1e-3 relative of it; DF K positive semidefinite for D = C C^T;
* the Cholesky ladder of a singular metric returns a finite factor.
"""

from types import SimpleNamespace

import numpy as np
import pytest

import mmd_ref
from asuka_amd.basis import build_ao_basis_cart
from asuka_amd.molecule import water_monomer


def _prim_basis(shells):
    """Single-primitive shells [(l, exponent, coefficient, centre)]."""
    nc = lambda l: (l + 1) * (l + 2) // 2
    ao = np.cumsum([0] + [nc(l) for l, *_ in shells])[:-1]
    return SimpleNamespace(
        shell_cxyz=np.array([c for *_, c in shells], dtype=np.float64).reshape(-1, 3),
        shell_prim_start=np.arange(len(shells), dtype=np.int32),
        shell_nprim=np.ones(len(shells), dtype=np.int32),
        shell_l=np.array([l for l, *_ in shells], dtype=np.int32),
        shell_ao_start=ao.astype(np.int32),
        prim_exp=np.array([e for _, e, *_ in shells], dtype=np.float64),
        prim_coef=np.array([c for _, _, c, _ in shells], dtype=np.float64))


def test_three_and_two_centre_against_mmd(oracle):
    rng = np.random.default_rng(3)
    A, B, C = rng.uniform(-1.5, 1.5, (3, 3))
    for la, lb, lp, lq in [(0, 0, 0, 0), (1, 0, 2, 1), (2, 1, 3, 0), (2, 2, 1, 3), (1, 1, 3, 3)]:
        a, b, p, q = rng.uniform(0.3, 3.0, 4)
        ca, cb, cp, cq = rng.uniform(0.5, 2.0, 4)
        ao = _prim_basis([(la, a, ca, A), (lb, b, cb, B)])
        aux = _prim_basis([(lp, p, cp, C), (lq, q, cq, B)])
        X, V = oracle.df_integrals(ao, aux, hp=True)
        nc = lambda l: (l + 1) * (l + 2) // 2
        na, nb, npp = nc(la), nc(lb), nc(lp)
        # (a b | p): X[P][mu][nu] with mu on shell a, nu on shell b
        ref = ca * cb * cp * mmd_ref.eri_prim_tile(la, lb, lp, 0, a, b, p, 0.0, A, B, C, C)
        ref = ref.reshape(na, nb, npp)
        got = X[:npp, :na, na:na + nb].transpose(1, 2, 0)
        assert np.abs(got - ref).max() < 1e-12 * max(1.0, np.abs(ref).max()), (la, lb, lp)
        # (p | q)
        refv = cp * cq * mmd_ref.eri_prim_tile(lp, 0, lq, 0, p, 0.0, q, 0.0, C, C, B, B)
        assert np.abs(V[:npp, npp:] - refv.reshape(npp, -1)).max() < 1e-12 * max(1.0, np.abs(refv).max())
        assert np.abs(V - V.T).max() < 1e-12 * np.abs(V).max()


def test_df_jk_water_against_exact(oracle):
    from asuka_amd.df import build_aux_basis_cart

    mol = water_monomer("cc-pVDZ")
    ao, _ = build_ao_basis_cart(mol)
    aux = build_aux_basis_cart(mol, None, "cc-pVDZ")
    X, V = oracle.df_integrals(ao, aux, hp=True)
    C = np.random.default_rng(4).standard_normal((ao.nao, 5))
    C, _ = np.linalg.qr(C)
    D = 2.0 * C @ C.T
    J, K = oracle.df_JK(X, V, D)
    octx = oracle.make_context(ao)
    J0, K0 = oracle.dense_JK(oracle.dense_eri(octx, hp=False), D)
    ej, ej0 = 0.5 * np.sum(D * J), 0.5 * np.sum(D * J0)
    assert ej <= ej0 + 1e-10
    assert (ej0 - ej) < 1e-3 * ej0
    assert np.abs(K - K0).max() < 0.05 * np.abs(K0).max()
    w = np.linalg.eigvalsh(K)
    assert w.min() > -1e-10 * w.max()


def test_cholesky_ladder_regularises_singular_metric(oracle):
    rng = np.random.default_rng(8)
    M = rng.standard_normal((6, 4))
    V = M @ M.T  # rank 4
    L = oracle.df_cholesky(V)
    assert np.all(np.isfinite(L))
    assert np.abs(L @ L.T - V).max() < 1e-8
