"""Pins the oracle's quartet screening (CPU, no GPU).

The oracle's shell-pair density maxima, its CSAM density mask
(hf/direct_jk.py:1473-1527, :1893-1909) on top of the Q-sorted Schwarz list
(cueri/tasks.py:49-115) and its per-class counter (screen_counts) are checked
against brute force on (H2O)2 cc-pVDZ: element-by-element block maxima, an
explicit double loop over every canonical quartet in the product form of
cueri/tasks.py build_tasks_screened, and the effect of the dropped quartets
on J and K, bounded by Cauchy-Schwarz.
"""

import numpy as np
import pytest

from asuka_amd.basis import build_ao_basis_cart
from asuka_amd.molecule import water_cluster

EPS_S = 1e-8
EPS_D = 1e-6


@pytest.fixture(scope="module")
def case(oracle):
    mol = water_cluster(2, 2, "cc-pVDZ")
    ao, _ = build_ao_basis_cart(mol)
    octx = oracle.make_context(ao)
    Q = oracle.schwarz(octx, hp=True)
    C = np.random.default_rng(31).standard_normal((ao.nao, 10))
    C, _ = np.linalg.qr(C)
    D = 2.0 * C @ C.T
    return ao, octx, Q, D


def _ncart(l):
    return (l + 1) * (l + 2) // 2


def _dsp_brute(ao, octx, D):
    out = np.zeros(octx.nsp)
    for sp in range(octx.nsp):
        A, B = int(octx.sp_A[sp]), int(octx.sp_B[sp])
        a0, b0 = int(ao.shell_ao_start[A]), int(ao.shell_ao_start[B])
        m = 0.0
        for i in range(_ncart(int(ao.shell_l[A]))):
            for j in range(_ncart(int(ao.shell_l[B]))):
                m = max(m, abs(float(D[a0 + i, b0 + j])), abs(float(D[b0 + j, a0 + i])))
        out[sp] = m
    return out


def _kept_brute(Q, dsp, eps_s, eps_d):
    kept = set()
    for ab in range(Q.size):
        for cd in range(ab + 1):
            qq = float(Q[ab]) * float(Q[cd])
            if qq > 0.0 and qq >= eps_s and qq * max(float(dsp[ab]), float(dsp[cd])) >= eps_d:
                kept.add((ab, cd))
    return kept


def test_shell_pair_dmax_is_the_block_maximum(oracle, case):
    ao, octx, Q, D = case
    assert np.array_equal(oracle.shell_pair_dmax(octx, D), _dsp_brute(ao, octx, D))


def test_density_screen_equals_brute_force(oracle, case):
    ao, octx, Q, D = case
    dsp = _dsp_brute(ao, octx, D)
    want = _kept_brute(Q, dsp, EPS_S, EPS_D)
    ab, cd = oracle.tasks_screened_sorted_q(Q, EPS_S)
    sab, scd = oracle.density_screen(octx, Q, D, ab, cd, EPS_D)
    got = set(zip(sab.tolist(), scd.tolist()))
    assert len(got) == sab.size  # no duplicates
    assert got == want
    # both cuts bite on this input
    ncanon = octx.nsp * (octx.nsp + 1) // 2
    assert ab.size < 0.95 * ncanon and sab.size < 0.95 * ab.size


def test_screen_counts_equal_brute_force(oracle, case):
    ao, octx, Q, D = case
    dsp = _dsp_brute(ao, octx, D)
    L = np.asarray(ao.shell_l)
    la, lb = L[octx.sp_A], L[octx.sp_B]
    cls = la * (la + 1) // 2 + lb
    neq = octx.sp_A != octx.sp_B
    order = np.argsort(-Q, kind="stable")
    rank = np.empty_like(order)
    rank[order] = np.arange(order.size)
    nprim = oracle.nonzero_prim_pairs(octx)
    for eps_d, dd in ((0.0, None), (EPS_D, dsp)):
        want = {}
        for ab, cd in _kept_brute(Q, dsp, EPS_S, eps_d):
            if cls[ab] != cls[cd]:
                bra, ket = (ab, cd) if cls[ab] > cls[cd] else (cd, ab)
            else:
                bra, ket = (ab, cd) if rank[ab] > rank[cd] else (cd, ab)
            key = (int(la[bra]), int(lb[bra]), int(la[ket]), int(lb[ket]))
            row = want.setdefault(key, np.zeros(9, np.int64))
            row[(int(neq[bra]) << 2) | (int(neq[ket]) << 1) | int(bra != ket)] += 1
            row[8] += nprim[bra] * nprim[ket]
        got = oracle.screen_counts(octx, Q, dd, EPS_S, eps_d)
        assert set(got) == set(want)
        for k in want:
            assert np.array_equal(got[k], want[k]), k


def test_dropped_quartets_bound_the_jk_change(oracle, case):
    """Screened minus unscreened J and K stay within the Cauchy-Schwarz bound
    of the dropped quartets: |(mn|ls)| <= Q_ab Q_cd, each quartet touching an
    element through at most 4 blocks of at most ncomp elements each (J: the
    density blocks of the pairs, Dsp; K: the cross blocks, bounded by max|D|)."""
    ao, octx, Q, D = case
    ab, cd = oracle.tasks_screened_sorted_q(Q, 0.0)
    J0, K0 = oracle.direct_jk_tasks(octx, D, ab, cd, hp=False)
    sab, scd = oracle.density_screen(octx, Q, D, *oracle.tasks_screened_sorted_q(Q, EPS_S), EPS_D)
    J1, K1 = oracle.direct_jk_tasks(octx, D, sab, scd, hp=False)
    kept = set(zip(sab.tolist(), scd.tolist()))
    dsp = oracle.shell_pair_dmax(octx, D)
    L = np.asarray(ao.shell_l)
    nsh = np.array([_ncart(int(L[a])) * _ncart(int(L[b])) for a, b in zip(octx.sp_A, octx.sp_B)])
    dmax = float(np.abs(D).max())
    bj = bk = 0.0
    for a, c in zip(ab.tolist(), cd.tolist()):
        if (a, c) in kept:
            continue
        qq = Q[a] * Q[c]
        n = max(nsh[a], nsh[c])
        bj += 4.0 * n * qq * max(dsp[a], dsp[c])
        bk += 4.0 * n * qq * dmax
    dJ = float(np.abs(J1 - J0).max())
    dK = float(np.abs(K1 - K0).max())
    assert 0.0 < dJ <= bj and 0.0 < dK <= bk
