"""Independent textbook McMurchie-Davidson integrals for primitive Cartesian
Gaussians (Helgaker, Jorgensen, Olsen, "Molecular Electronic-Structure Theory",
ch. 9), used only to pin the oracle on small inputs.

Nothing here follows the reference's code paths: Boys values come from the
all-positive series F_M(T) = e^-T sum_k (2T)^k / ((2M+1)(2M+3)...(2M+2k+1))
followed by downward recursion, and integrals from Hermite expansions.
"""

from __future__ import annotations

import math

import numpy as np


def boys_array(m_max: int, T: float) -> np.ndarray:
    M = m_max + 8
    term = 1.0 / (2 * M + 1)
    s = term
    k = 0
    while True:
        k += 1
        term *= 2.0 * T / (2 * M + 2 * k + 1)
        s += term
        if term < 1e-18 * s and k > T:
            break
    F = np.zeros(M + 1)
    e = math.exp(-T)
    F[M] = e * s
    for m in range(M - 1, -1, -1):
        F[m] = (2.0 * T * F[m + 1] + e) / (2 * m + 1)
    return F[: m_max + 1]


def hermite_E(i: int, j: int, t: int, Qx: float, a: float, b: float) -> float:
    """E_t^{ij} with Qx = A - B (includes exp(-mu Qx^2))."""
    p = a + b
    q = a * b / p
    if t < 0 or t > i + j:
        return 0.0
    if i == j == t == 0:
        return math.exp(-q * Qx * Qx)
    if j == 0:
        return (1.0 / (2 * p)) * hermite_E(i - 1, j, t - 1, Qx, a, b) - (q * Qx / a) * hermite_E(
            i - 1, j, t, Qx, a, b) + (t + 1) * hermite_E(i - 1, j, t + 1, Qx, a, b)
    return (1.0 / (2 * p)) * hermite_E(i, j - 1, t - 1, Qx, a, b) + (q * Qx / b) * hermite_E(
        i, j - 1, t, Qx, a, b) + (t + 1) * hermite_E(i, j - 1, t + 1, Qx, a, b)


def R_table(L: int, alpha: float, PC: np.ndarray) -> dict:
    """R^0_{tuv} for t+u+v <= L (Hermite Coulomb integrals)."""
    T = alpha * float(PC @ PC)
    F = boys_array(L, T)
    memo = {}

    def R(n, t, u, v):
        key = (n, t, u, v)
        if key in memo:
            return memo[key]
        if t < 0 or u < 0 or v < 0:
            val = 0.0
        elif t == u == v == 0:
            val = (-2.0 * alpha) ** n * F[n]
        elif t > 0:
            val = (t - 1) * R(n + 1, t - 2, u, v) + PC[0] * R(n + 1, t - 1, u, v)
        elif u > 0:
            val = (u - 1) * R(n + 1, t, u - 2, v) + PC[1] * R(n + 1, t, u - 1, v)
        else:
            val = (v - 1) * R(n + 1, t, u, v - 2) + PC[2] * R(n + 1, t, u, v - 1)
        memo[key] = val
        return val

    return R


def cart(l: int):
    return [(x, y, l - x - y) for x in range(l, -1, -1) for y in range(l - x, -1, -1)]


def eri_prim_tile(la, lb, lc, ld, a, b, c, d, A, B, C, D) -> np.ndarray:
    """(ab|cd) for unit-coefficient primitives, tile (nA*nB, nC*nD)."""
    A, B, C, D = (np.asarray(v, dtype=float) for v in (A, B, C, D))
    p, q = a + b, c + d
    P = (a * A + b * B) / p
    Q = (c * C + d * D) / q
    alpha = p * q / (p + q)
    R = R_table(la + lb + lc + ld, alpha, P - Q)
    pref = 2.0 * math.pi ** 2.5 / (p * q * math.sqrt(p + q))
    ca, cb, cc, cd = cart(la), cart(lb), cart(lc), cart(ld)
    out = np.zeros((len(ca) * len(cb), len(cc) * len(cd)))
    for ia, ea in enumerate(ca):
        for ib, eb in enumerate(cb):
            Eab = [[hermite_E(ea[k], eb[k], t, A[k] - B[k], a, b) for t in range(ea[k] + eb[k] + 1)] for k in range(3)]
            for ic, ec in enumerate(cc):
                for id_, ed in enumerate(cd):
                    Ecd = [[hermite_E(ec[k], ed[k], t, C[k] - D[k], c, d) for t in range(ec[k] + ed[k] + 1)]
                           for k in range(3)]
                    s = 0.0
                    for t, ex in enumerate(Eab[0]):
                        for u, ey in enumerate(Eab[1]):
                            for v, ez in enumerate(Eab[2]):
                                e1 = ex * ey * ez
                                if e1 == 0.0:
                                    continue
                                for tau, fx in enumerate(Ecd[0]):
                                    for nu, fy in enumerate(Ecd[1]):
                                        for phi, fz in enumerate(Ecd[2]):
                                            s += e1 * fx * fy * fz * (-1) ** (tau + nu + phi) * R(
                                                0, t + tau, u + nu, v + phi)
                    out[ia * len(cb) + ib, ic * len(cd) + id_] = pref * s
    return out


def overlap_prim(la, lb, a, b, A, B) -> np.ndarray:
    A, B = np.asarray(A, float), np.asarray(B, float)
    p = a + b
    out = np.zeros((len(cart(la)), len(cart(lb))))
    for i, ea in enumerate(cart(la)):
        for j, eb in enumerate(cart(lb)):
            v = (math.pi / p) ** 1.5
            for k in range(3):
                v *= hermite_E(ea[k], eb[k], 0, A[k] - B[k], a, b)
            out[i, j] = v
    return out


def nuclear_prim(la, lb, a, b, A, B, Cn, Z) -> np.ndarray:
    A, B, Cn = np.asarray(A, float), np.asarray(B, float), np.asarray(Cn, float)
    p = a + b
    P = (a * A + b * B) / p
    R = R_table(la + lb, p, P - Cn)
    out = np.zeros((len(cart(la)), len(cart(lb))))
    for i, ea in enumerate(cart(la)):
        for j, eb in enumerate(cart(lb)):
            s = 0.0
            for t in range(ea[0] + eb[0] + 1):
                for u in range(ea[1] + eb[1] + 1):
                    for v in range(ea[2] + eb[2] + 1):
                        s += (hermite_E(ea[0], eb[0], t, A[0] - B[0], a, b) * hermite_E(ea[1], eb[1], u, A[1] - B[1], a, b)
                              * hermite_E(ea[2], eb[2], v, A[2] - B[2], a, b) * R(0, t, u, v))
            out[i, j] = -Z * 2.0 * math.pi / p * s
    return out
