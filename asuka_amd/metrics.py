"""Algorithmic work of a direct J/K build: the fixed counting convention "EFM-v1".

FLOPs per screened shell quartet q of class (la, lb, lc, ld), with
L = la+lb+lc+ld, n = L//2 + 1 roots, nmax = la+lb, mmax = lc+ld,
ncomp = prod ncart(l), N1 = prod (l+1), K = primitive quartets with a nonzero
coefficient product:

  FLOP_eri(q) = K * [24 + n * (24 + 15 (nmax+1)(mmax+1) + 6 N1 + 4 ncomp)]
  FLOP_dig(q) = 2 ncomp (jt + kt),  jt = 1 + [bra pair != ket pair],
                kt = 1 + [C!=D] + [A!=B] + [A!=B and C!=D]

(24 = primitive-quartet setup; per root 24 for the recursion coefficients,
5 per 2-D VRR entry per axis, 2 per 1-D transfer entry per axis and 4 per
component for the product and accumulation; Boys/root evaluation is excluded.
The digestion term is twice the reference's per-density inner multiplies,
asuka/hf/direct_jk_cost.py:196-203.)

Algorithmic bytes follow the reference's warp-contraction model
(asuka/hf/direct_jk_cost.py:219-266): 8 B per density load
(nAB nCD (jt + kt) per quartet) plus 16 B per output atomic (a read-modify-write):
J: nAB (1+[A!=B]) + [bk] nCD (1+[C!=D]); K: (1+[bk]) (nA nC + [C!=D] nA nD +
[A!=B] nB nC + [A!=B][C!=D] nB nD).

Inputs are the per-class-pair symmetry histograms the kernels count
(bin = (A!=B)<<2 | (C!=D)<<1 | bk).

For the nuclear gradient of the two-electron energy, per quartet

  FLOP_grad(q) = K * [24 + n' (24 + 15 (nmax+2)(mmax+2) + 6 N1' + 12 * 4 ncomp)]

with n' = (L+1)//2 + 1 roots and N1' = prod (l+2): the recursions run one
order higher and each component contracts 12 (centre, xyz) derivative
products.
"""

from __future__ import annotations

from dataclasses import dataclass


def ncart(l: int) -> int:
    return (l + 1) * (l + 2) // 2


def flop_eri(la: int, lb: int, lc: int, ld: int, K: int) -> int:
    la, lb, lc, ld, K = int(la), int(lb), int(lc), int(ld), int(K)
    L = la + lb + lc + ld
    n = L // 2 + 1
    ncomp = ncart(la) * ncart(lb) * ncart(lc) * ncart(ld)
    N1 = (la + 1) * (lb + 1) * (lc + 1) * (ld + 1)
    per_root = 24 + 15 * (la + lb + 1) * (lc + ld + 1) + 6 * N1 + 4 * ncomp
    return int(K) * (24 + n * per_root)


def flop_grad(la: int, lb: int, lc: int, ld: int, K: int) -> int:
    la, lb, lc, ld, K = int(la), int(lb), int(lc), int(ld), int(K)
    L = la + lb + lc + ld
    n = (L + 1) // 2 + 1
    ncomp = ncart(la) * ncart(lb) * ncart(lc) * ncart(ld)
    N1 = (la + 2) * (lb + 2) * (lc + 2) * (ld + 2)
    per_root = 24 + 15 * (la + lb + 2) * (lc + ld + 2) + 6 * N1 + 12 * 4 * ncomp
    return int(K) * (24 + n * per_root)


def _terms(bin_: int) -> tuple[int, int]:
    ab_neq, cd_neq, bk = (bin_ >> 2) & 1, (bin_ >> 1) & 1, bin_ & 1
    return 1 + bk, 1 + cd_neq + ab_neq + (ab_neq & cd_neq)


def flop_dig(la: int, lb: int, lc: int, ld: int, hist) -> int:
    la, lb, lc, ld = int(la), int(lb), int(lc), int(ld)
    ncomp = ncart(la) * ncart(lb) * ncart(lc) * ncart(ld)
    return sum(2 * ncomp * sum(_terms(b)) * int(c) for b, c in enumerate(hist))


def bytes_alg(la: int, lb: int, lc: int, ld: int, hist) -> int:
    la, lb, lc, ld = int(la), int(lb), int(lc), int(ld)
    nA, nB, nC, nD = ncart(la), ncart(lb), ncart(lc), ncart(ld)
    nAB, nCD = nA * nB, nC * nD
    total = 0
    for b, c in enumerate(hist):
        if not c:
            continue
        ab_neq, cd_neq, bk = (b >> 2) & 1, (b >> 1) & 1, b & 1
        jt, kt = _terms(b)
        loads = nAB * nCD * (jt + kt)
        atom_j = nAB * (1 + ab_neq) + bk * nCD * (1 + cd_neq)
        atom_k = (1 + bk) * (nA * nC + cd_neq * nA * nD + ab_neq * nB * nC + ab_neq * cd_neq * nB * nD)
        total += int(c) * (8 * loads + 16 * (atom_j + atom_k))
    return total


@dataclass
class BuildWork:
    quartets: int = 0
    prim_quartets: int = 0
    prim_roots: int = 0
    flop_eri: int = 0
    flop_dig: int = 0
    bytes_alg: int = 0

    @property
    def flops(self) -> int:
        return self.flop_eri + self.flop_dig


def build_work(class_rows: list[tuple[tuple[int, int, int, int], list[int], int]]) -> BuildWork:
    """class_rows: ((la, lb, lc, ld), histogram[8], primitive quartets K_total).

    FLOP_eri is linear in K, so the class total uses K_total directly."""
    w = BuildWork()
    for (la, lb, lc, ld), hist, ksum in class_rows:
        nq = int(sum(hist))
        w.quartets += nq
        w.prim_quartets += int(ksum)
        n = (la + lb + lc + ld) // 2 + 1
        w.prim_roots += int(ksum) * n
        w.flop_eri += flop_eri(la, lb, lc, ld, int(ksum))
        w.flop_dig += flop_dig(la, lb, lc, ld, hist)
        w.bytes_alg += bytes_alg(la, lb, lc, ld, hist)
    return w


# MI355X published peaks
PEAK_FP64_TFLOPS = 78.6
PEAK_FP32_TFLOPS = 157.3
PEAK_HBM_GBS = 8000.0


def roofline_fraction(f64: float, f32: float, seconds: float) -> float:
    """(F64/P64 + F32/P32) / t."""
    return (f64 / (PEAK_FP64_TFLOPS * 1e12) + f32 / (PEAK_FP32_TFLOPS * 1e12)) / seconds


__all__ = ["BuildWork", "PEAK_FP32_TFLOPS", "PEAK_FP64_TFLOPS", "PEAK_HBM_GBS", "build_work", "bytes_alg",
           "flop_dig", "flop_eri", "flop_grad", "roofline_fraction"]
