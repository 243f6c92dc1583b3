"""Write ``rys_golden.npz``: high-precision Boys values and Rys roots/weights.

Independent of the product's table generator: every quantity here comes from
direct numerical quadrature with mpmath (``mp.quad``) of the Rys weight
exp(-T t^2) on t in [0, 1], never from Boys-moment recurrences or 1F1:

* Boys F_m(T) = int_0^1 t^(2m) exp(-T t^2) dt = gamma(m+1/2, T) / (2 T^(m+1/2))
  (lower incomplete gamma, ``mp.gammainc``; 1/(2m+1) at T = 0);
* Rys rule of order n: discretised Stieltjes procedure on that weight
  (recurrence coefficients from quadrature inner products), then the
  eigen-decomposition of the Jacobi matrix; roots are x = t^2.

Run: ``python tests/golden/make_rys_golden.py`` (a few minutes).
"""

from __future__ import annotations

import os
from multiprocessing import Pool

import mpmath as mp
import numpy as np

DPS = 40
M_MAX = 16
T_VALUES = [0.0, 1e-14, 1e-9, 1e-5, 0.01, 0.1, 0.5, 1.0, 1.999, 2.0, 2.001, 3.3, 4.99, 5.0, 5.01, 7.7, 10.0,
            13.1, 18.0, 25.0, 33.3, 39.99, 40.01, 47.5, 55.0, 63.0, 71.0, 79.9, 88.0, 95.0, 120.0, 300.0, 2000.0]
N_MAX = 7


def _inner(f, T):
    g = lambda t: f(t * t) * mp.exp(-T * t * t)
    # breakpoints follow the width 1/sqrt(T) of the peaked integrand
    h = mp.mpf(1) / 4 if T < 16 else 1 / (2 * mp.sqrt(T))
    pts = [mp.mpf(0)]
    while pts[-1] + h < 1:
        pts.append(pts[-1] + h)
    pts.append(mp.mpf(1))
    return mp.quad(g, pts)


def boys(m, T):
    mp.mp.dps = DPS
    T = mp.mpf(T)
    if T == 0:
        return mp.mpf(1) / (2 * m + 1)
    a = m + mp.mpf(1) / 2
    return mp.gammainc(a, 0, T) / (2 * T ** a)


def rys(n, T):
    # the monomial-basis Stieltjes polynomials cancel like T^n at large T
    mp.mp.dps = DPS if T < 100 else DPS + 60
    T = mp.mpf(T)
    # polynomials in x as coefficient lists (ascending)
    def ev(c, x):
        s = mp.mpf(0)
        for a in reversed(c):
            s = s * x + a
        return s

    alpha, beta = [], []
    p_prev, p_cur = [mp.mpf(0)], [mp.mpf(1)]
    norm_prev = None
    for k in range(n):
        nk = _inner(lambda x: ev(p_cur, x) ** 2, T)
        ak = _inner(lambda x: x * ev(p_cur, x) ** 2, T) / nk
        alpha.append(ak)
        beta.append(nk / norm_prev if norm_prev is not None else nk)
        # p_{k+1} = (x - a_k) p_k - b_k p_{k-1}
        nxt = [mp.mpf(0)] * (len(p_cur) + 1)
        for i, c in enumerate(p_cur):
            nxt[i + 1] += c
            nxt[i] -= ak * c
        if k > 0:
            for i, c in enumerate(p_prev):
                nxt[i] -= beta[k] * c
        p_prev, p_cur, norm_prev = p_cur, nxt, nk
    J = mp.matrix(n, n)
    for i in range(n):
        J[i, i] = alpha[i]
        if i + 1 < n:
            J[i, i + 1] = J[i + 1, i] = mp.sqrt(beta[i + 1])
    E, Q = mp.eigsy(J)
    order = sorted(range(n), key=lambda i: E[i])
    mu0 = beta[0]
    return [float(E[i]) for i in order], [float(mu0 * Q[0, i] ** 2) for i in order]


def _boys_row(T):
    return [float(boys(m, T)) for m in range(M_MAX + 1)]


def _rys_row(args):
    n, T = args
    return rys(n, T)


def main():
    with Pool(min(8, os.cpu_count() or 1)) as pool:
        F = np.asarray(pool.map(_boys_row, T_VALUES))
        jobs = [(n, T) for n in range(1, N_MAX + 1) for T in T_VALUES]
        res = pool.map(_rys_row, jobs)
    X = np.zeros((N_MAX + 1, len(T_VALUES), N_MAX))
    W = np.zeros_like(X)
    for (n, T), (r, w) in zip(jobs, res):
        i = T_VALUES.index(T)
        X[n, i, :n] = r
        W[n, i, :n] = w
    out = os.path.join(os.path.dirname(os.path.abspath(__file__)), "rys_golden.npz")
    np.savez(out, T=np.asarray(T_VALUES), boys=F, x=X, w=W)
    print("wrote", out)


if __name__ == "__main__":
    main()
