"""Generate the Rys root/weight interpolation table (``rys_table.h``).

Rys quadrature of order n for the weight exp(-T x) x^(-1/2) on x in [0, 1]
(x = t^2) has moments mu_k = F_k(T), the Boys functions. For each n = 1..NMAX
the roots x_u(T) and weights w_u(T) are

* for T < TMAX[n]: piecewise polynomials on intervals of width H, degree DEG,
  fitted by Chebyshev interpolation to values computed with mpmath at 50
  digits (Boys moments from 1F1, then Stieltjes recurrence coefficients and a
  Jacobi-matrix eigensolve), stored in the power basis of s = 2(T-T_lo)/H - 1;
* for T >= TMAX[n]: the large-T limit x_u = h_u^2/T, w_u = W_u/sqrt(T) with
  (h_u, W_u) the positive nodes and weights of 2n-point Gauss-Hermite
  quadrature. TMAX[n] is the first interval edge where that limit agrees with
  the 50-digit values to better than 1e-15 relative.

Run ``python gen_rys_table.py`` to rewrite ``rys_table.h`` next to this file.
The fit is checked at off-node points; the script fails if any relative error
exceeds 5e-15.
"""

from __future__ import annotations

import os
import sys
from multiprocessing import Pool

import mpmath as mp
import numpy as np

NMAX = 7
H = 2.0
DEG = 13
DPS = 50


def boys_mp(m: int, T) -> mp.mpf:
    T = mp.mpf(T)
    return mp.hyp1f1(m + mp.mpf(1) / 2, m + mp.mpf(3) / 2, -T) / (2 * m + 1)


def rys_mp(n: int, T) -> tuple[list, list]:
    """n-point Rys roots (x = t^2, ascending) and weights at T, in mpmath."""
    mp.mp.dps = DPS
    mu = [boys_mp(k, T) for k in range(2 * n)]
    alpha = [mp.mpf(0)] * n
    beta = [mp.mpf(0)] * n
    p_prev = [mp.mpf(0)] * (n + 1)
    p_cur = [mp.mpf(0)] * (n + 1)
    p_cur[0] = mp.mpf(1)
    h = mu[0]
    alpha[0] = mu[1] / mu[0]

    def inner(a, shift):
        s = mp.mpf(0)
        for i in range(n + 1):
            if a[i] == 0:
                continue
            for j in range(n + 1):
                if i + j + shift < 2 * n:
                    s += a[i] * a[j] * mu[i + j + shift]
        return s

    for k in range(n - 1):
        pn = [mp.mpf(0)] * (n + 1)
        for i in range(k + 1):
            pn[i + 1] += p_cur[i]
            pn[i] -= alpha[k] * p_cur[i]
        if k >= 1:
            for i in range(k):
                pn[i] -= beta[k] * p_prev[i]
        hn = inner(pn, 0)
        beta[k + 1] = hn / h
        alpha[k + 1] = inner(pn, 1) / hn
        p_prev, p_cur, h = p_cur, pn, hn
    J = mp.matrix(n, n)
    for i in range(n):
        J[i, i] = alpha[i]
        if i + 1 < n:
            J[i, i + 1] = J[i + 1, i] = mp.sqrt(beta[i + 1])
    E, Q = mp.eigsy(J)
    order = sorted(range(n), key=lambda i: E[i])
    return [E[i] for i in order], [mu[0] * Q[0, i] ** 2 for i in order]


_HERM = {}


def hermite_positive(n: int) -> tuple[np.ndarray, np.ndarray]:
    """Positive nodes h_u (ascending) and weights W_u of the 2n-point
    Gauss-Hermite rule, from mpmath at DPS digits; returns (h_u^2, W_u)."""
    if n not in _HERM:
        mp.mp.dps = DPS
        m = 2 * n
        # Jacobi matrix of the Hermite weight exp(-x^2): off-diagonal sqrt(k/2)
        J = mp.matrix(m, m)
        for k in range(1, m):
            J[k - 1, k] = J[k, k - 1] = mp.sqrt(mp.mpf(k) / 2)
        E, Q = mp.eigsy(J)
        pairs = sorted((E[i], mp.sqrt(mp.pi) * Q[0, i] ** 2) for i in range(m) if E[i] > 0)
        _HERM[n] = (np.array([float(h * h) for h, _ in pairs]), np.array([float(W) for _, W in pairs]))
    return _HERM[n]


def hermite_limit(n: int, T: float) -> tuple[np.ndarray, np.ndarray]:
    hx, hw = hermite_positive(n)
    return hx / T, hw / np.sqrt(T)


def _eval_point(args):
    n, T = args
    r, w = rys_mp(n, T)
    return [float(v) for v in r] + [float(v) for v in w]


def _pick_tmax(n: int, pool: Pool) -> float:
    T = 20.0
    while True:
        pts = [T, T + 0.37 * H, T + 0.81 * H, T + 5.0, T + 20.0]
        vals = pool.map(_eval_point, [(n, t) for t in pts])
        ok = True
        for t, v in zip(pts, vals):
            hr, hw = hermite_limit(n, t)
            ref = np.asarray(v)
            got = np.concatenate([hr, hw])
            if np.max(np.abs(got - ref) / np.abs(ref)) > 1e-15:
                ok = False
                break
        if ok:
            return T
        T += H
        if T > 400:
            raise SystemExit(f"no asymptotic cutoff found for n={n}")


def main(out_path: str) -> None:
    m = DEG + 1
    kk = np.arange(m)
    s_nodes = np.cos(np.pi * (kk + 0.5) / m)
    tables = {}
    tmax = {}
    with Pool(min(8, os.cpu_count() or 1)) as pool:
        for n in range(1, NMAX + 1):
            tmax[n] = _pick_tmax(n, pool)
            nint = int(round(tmax[n] / H))
            jobs = [(n, i * H + H * (s + 1.0) / 2.0) for i in range(nint) for s in s_nodes]
            vals = np.asarray(pool.map(_eval_point, jobs)).reshape(nint, m, 2 * n)
            coef = np.empty((nint, 2 * n, m))
            for i in range(nint):
                ch = np.polynomial.chebyshev.chebfit(s_nodes, vals[i], DEG)  # (m, 2n)
                for f in range(2 * n):
                    coef[i, f] = np.polynomial.chebyshev.cheb2poly(ch[:, f])
            # Off-node check, evaluated exactly as the kernels do (Horner in s).
            checks = [(n, i * H + H * (s + 1.0) / 2.0) for i in range(nint) for s in (-0.93, -0.41, 0.17, 0.66, 0.995)]
            ref = np.asarray(pool.map(_eval_point, checks))
            worst = 0.0
            for (nn, T), rv in zip(checks, ref):
                i = min(int(T / H), nint - 1)
                s = 2.0 * (T - i * H) / H - 1.0
                got = np.zeros(2 * n)
                for f in range(2 * n):
                    acc = coef[i, f, DEG]
                    for d in range(DEG - 1, -1, -1):
                        acc = acc * s + coef[i, f, d]
                    got[f] = acc
                worst = max(worst, float(np.max(np.abs(got - rv) / np.abs(rv))))
            print(f"n={n}: TMAX={tmax[n]:.0f} intervals={nint} worst_rel_err={worst:.2e}", flush=True)
            if worst > 5e-15:
                raise SystemExit(f"fit error too large for n={n}: {worst}")
            tables[n] = coef

    herm = {n: hermite_positive(n) for n in range(1, NMAX + 1)}

    lines = []
    lines.append("// Generated by gen_rys_table.py -- do not edit by hand.")
    lines.append("// Rys roots x_u(T) (x = t^2) and weights w_u(T) for n = 1..RYS_NMAX roots:")
    lines.append("// piecewise polynomials of degree RYS_DEG in s = 2(T - i*RYS_H)/RYS_H - 1 on")
    lines.append("// interval i for T < rys_tmax[n]; beyond it x_u = hx_u/T, w_u = hw_u/sqrt(T)")
    lines.append("// (2n-point Gauss-Hermite limit). Layout per n: [interval][2n functions][RYS_DEG+1],")
    lines.append("// functions 0..n-1 are roots (ascending), n..2n-1 the matching weights.")
    lines.append("// Included once per address space and coefficient type: define RYS_TABLE_QUAL")
    lines.append("// (e.g. `static const` or `static const __device__`) and RYS_COEF_T (double, or")
    lines.append("// float for an FP32 copy of the coefficients) and wrap the include in a namespace.")
    lines.append(f"#define RYS_NMAX {NMAX}")
    lines.append(f"#define RYS_DEG {DEG}")
    lines.append(f"#define RYS_H {H!r}")
    lines.append(f"#define RYS_INV_H {1.0 / H!r}")
    nints = [0] + [int(round(tmax[n] / H)) for n in range(1, NMAX + 1)]
    offs = [0]
    for n in range(1, NMAX + 1):
        offs.append(offs[-1] + nints[n] * 2 * n * (DEG + 1))
    # the same metadata as macro lists, for compile-time use (constexpr copies)
    lines.append("#define RYS_NINT_LIST " + ", ".join(str(v) for v in nints))
    lines.append("#define RYS_TMAX_LIST 0.0, " + ", ".join(f"{tmax[n]!r}" for n in range(1, NMAX + 1)))
    lines.append("#define RYS_COEF_OFF_LIST 0, " + ", ".join(str(v) for v in offs[:-1]))
    lines.append("RYS_TABLE_QUAL int rys_nint[RYS_NMAX + 1] = {" + ", ".join(str(v) for v in nints) + "};")
    lines.append("RYS_TABLE_QUAL double rys_tmax[RYS_NMAX + 1] = {0.0, " +
                 ", ".join(f"{tmax[n]!r}" for n in range(1, NMAX + 1)) + "};")
    # rys_coef_off[n] = start of the n-root block (index 0 unused).
    lines.append("RYS_TABLE_QUAL int rys_coef_off[RYS_NMAX + 1] = {0, " + ", ".join(str(v) for v in offs[:-1]) + "};")
    hx_flat, hw_flat, hoff = [], [], [0]
    for n in range(1, NMAX + 1):
        hx_flat.extend(herm[n][0].tolist())
        hw_flat.extend(herm[n][1].tolist())
    for n in range(1, NMAX + 1):
        hoff.append(hoff[-1] + n)
    lines.append("RYS_TABLE_QUAL int rys_herm_off[RYS_NMAX + 1] = {0, " + ", ".join(str(hoff[n - 1]) for n in range(1, NMAX + 1)) + "};")
    lines.append(f"RYS_TABLE_QUAL double rys_herm_x[{len(hx_flat)}] = {{" + ", ".join(repr(float(v)) for v in hx_flat) + "};")
    lines.append(f"RYS_TABLE_QUAL double rys_herm_w[{len(hw_flat)}] = {{" + ", ".join(repr(float(v)) for v in hw_flat) + "};")
    flat = []
    for n in range(1, NMAX + 1):
        flat.extend(tables[n].ravel().tolist())
    lines.append(f"#define RYS_COEF_TOTAL {len(flat)}")
    lines.append(f"RYS_TABLE_QUAL RYS_COEF_T rys_coef[{len(flat)}] = {{")
    for i in range(0, len(flat), 4):
        lines.append("  " + ", ".join(repr(float(v)) for v in flat[i:i + 4]) + ",")
    lines.append("};")
    with open(out_path, "w") as f:
        f.write("\n".join(lines) + "\n")
    print("wrote", out_path, "coefficients:", len(flat))


if __name__ == "__main__":
    here = os.path.dirname(os.path.abspath(__file__))
    main(sys.argv[1] if len(sys.argv) > 1 else os.path.join(here, "rys_table.h"))
