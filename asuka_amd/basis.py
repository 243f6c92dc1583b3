"""Cartesian AO basis: vendored basis data, parsing and packing.

The packed layout is a structure of arrays with one shell per contraction column,
Cartesian components ordered lx descending then ly descending, primitive
coefficients that already carry the primitive and contraction normalisation.
It follows the reference packer:

- shell order per element: grouped by l, stable within l
  (frontend/basis_bse.py:79-91 ``_reorder_shells_like_pyscf``);
- one shell per contraction column, every primitive of the block kept
  (frontend/basis_packer.py:183-251 ``pack_cart_basis``);
- primitive norms (integrals/gto_cart.py:42-66 ``primitive_norm_cart_like_pyscf``);
- contraction normalisation to a self-overlap of 1 (l <= 1) or 4*pi/(2l+1)
  (l >= 2) for the (l,0,0) component (frontend/basis_packer.py:35-93);
- Cartesian order (cueri/cart.py:28-54).

Basis data: the correlation-consistent sets cc-pVDZ and cc-pVTZ for H and O
(Dunning, J. Chem. Phys. 90, 1007 (1989)), written in the explicit
``{element: [[l, [exp, c1, c2, ...], ...], ...]}`` form that the reference's
``parse_pyscf_basis_dict`` (frontend/basis_packer.py:142-180) accepts. General
contractions are kept as multi-column blocks, so uncontracted functions appear
as unit columns over the block's exponents, as in the Basis Set Exchange JSON
that the reference loads through ``basis_set_exchange``.
"""

from __future__ import annotations

import math
import re
from dataclasses import dataclass
from typing import Any

import numpy as np

ANGSTROM_TO_BOHR = 1.8897259886  # frontend/molecule.py:10

# ---------------------------------------------------------------------------
# Vendored basis data
# ---------------------------------------------------------------------------

_H_DZ_S = [
    [13.01, 0.019685, 0.0],
    [1.962, 0.137977, 0.0],
    [0.4446, 0.478148, 0.0],
    [0.122, 0.50124, 1.0],
]
_O_DZ_S = [
    [11720.0, 0.00071, -0.00016, 0.0],
    [1759.0, 0.00547, -0.001263, 0.0],
    [400.8, 0.027837, -0.006267, 0.0],
    [113.7, 0.1048, -0.025716, 0.0],
    [37.03, 0.283062, -0.070924, 0.0],
    [13.27, 0.448719, -0.165411, 0.0],
    [5.025, 0.270952, -0.116955, 0.0],
    [1.013, 0.015458, 0.557368, 0.0],
    [0.3023, -0.002585, 0.572759, 1.0],
]
_O_DZ_P = [
    [17.7, 0.043018, 0.0],
    [3.854, 0.228913, 0.0],
    [1.046, 0.508728, 0.0],
    [0.2753, 0.460531, 1.0],
]

CC_PVDZ: dict[str, list] = {
    "H": [
        [0, *_H_DZ_S],
        [1, [0.727, 1.0]],
    ],
    "O": [
        [0, *_O_DZ_S],
        [1, *_O_DZ_P],
        [2, [1.185, 1.0]],
    ],
}

_H_TZ_S = [
    [33.87, 0.006068, 0.0, 0.0],
    [5.095, 0.045308, 0.0, 0.0],
    [1.159, 0.202822, 0.0, 0.0],
    [0.3258, 0.503903, 1.0, 0.0],
    [0.1027, 0.383421, 0.0, 1.0],
]
_O_TZ_S = [
    [15330.0, 0.000508, -0.000115, 0.0, 0.0],
    [2299.0, 0.003929, -0.000895, 0.0, 0.0],
    [522.4, 0.020243, -0.004636, 0.0, 0.0],
    [147.3, 0.079181, -0.018724, 0.0, 0.0],
    [47.55, 0.230687, -0.058463, 0.0, 0.0],
    [16.76, 0.433118, -0.136463, 0.0, 0.0],
    [6.207, 0.35026, -0.17574, 0.0, 0.0],
    [1.752, 0.042728, 0.160934, 1.0, 0.0],
    [0.6882, -0.008154, 0.603418, 0.0, 0.0],
    [0.2384, 0.002381, 0.378765, 0.0, 1.0],
]
_O_TZ_P = [
    [34.46, 0.015928, 0.0, 0.0],
    [7.749, 0.09974, 0.0, 0.0],
    [2.28, 0.310492, 0.0, 0.0],
    [0.7156, 0.491026, 1.0, 0.0],
    [0.214, 0.336337, 0.0, 1.0],
]

CC_PVTZ: dict[str, list] = {
    "H": [
        [0, *_H_TZ_S],
        [1, [1.407, 1.0, 0.0], [0.388, 0.0, 1.0]],
        [2, [1.057, 1.0]],
    ],
    "O": [
        [0, *_O_TZ_S],
        [1, *_O_TZ_P],
        [2, [2.314, 1.0, 0.0], [0.645, 0.0, 1.0]],
        [3, [1.428, 1.0]],
    ],
}

_NAMED = {"cc-pvdz": CC_PVDZ, "cc-pvtz": CC_PVTZ}

_ATOMIC_NUMBER = {"H": 1, "He": 2, "Li": 3, "Be": 4, "B": 5, "C": 6, "N": 7, "O": 8, "F": 9, "Ne": 10}


def atomic_number(sym: str) -> int:
    """Nuclear charge of an element symbol (chem/periodic_table.py)."""
    key = str(sym).strip().capitalize()
    m = re.match(r"^([A-Za-z]{1,2})", key)
    key = m.group(1).capitalize() if m else key
    if key not in _ATOMIC_NUMBER:
        raise KeyError(f"unknown element {sym!r}")
    return _ATOMIC_NUMBER[key]


def named_basis(name: str) -> dict[str, list]:
    """Return a vendored explicit basis dict by (case-insensitive) name."""
    key = str(name).strip().lower()
    if key not in _NAMED:
        raise KeyError(f"basis {name!r} is not vendored (available: {sorted(_NAMED)})")
    return _NAMED[key]


# ---------------------------------------------------------------------------
# Cartesian conventions
# ---------------------------------------------------------------------------


def ncart(l: int) -> int:
    if l < 0:
        raise ValueError("l must be >= 0")
    return (l + 1) * (l + 2) // 2


def cartesian_components(l: int) -> tuple[tuple[int, int, int], ...]:
    """(lx, ly, lz) in lx-descending then ly-descending order (cueri/cart.py:28-54)."""
    out = []
    for lx in range(l, -1, -1):
        for ly in range(l - lx, -1, -1):
            out.append((lx, ly, l - lx - ly))
    return tuple(out)


def _double_factorial(n: int) -> int:
    out = 1
    for k in range(int(n), 0, -2):
        out *= k
    return out


def primitive_norm(l: int, exps: np.ndarray) -> np.ndarray:
    """Primitive coefficient scale (integrals/gto_cart.py:42-66).

    l <= 1: (2a/pi)^(3/4) (4a)^(l/2); l >= 2: the radial norm
    1/sqrt(int_0^inf r^(2l+2) exp(-2a r^2) dr).
    """
    exps = np.asarray(exps, dtype=np.float64)
    if l <= 1:
        return (2.0 * exps / math.pi) ** 0.75 * (4.0 * exps) ** (0.5 * l)
    n1 = 0.5 * float(2 * l + 3)
    radial = (math.gamma(n1) / 2.0) / np.power(2.0 * exps, n1)
    return 1.0 / np.sqrt(radial)


def normalize_contraction(l: int, exps: np.ndarray, prim_coef: np.ndarray) -> np.ndarray:
    """Scale one contraction column to its (l,0,0) self-overlap target
    (frontend/basis_packer.py:50-93)."""
    exps = np.asarray(exps, dtype=np.float64)
    c = np.asarray(prim_coef, dtype=np.float64)
    alpha = exps[:, None] + exps[None, :]
    c_l = float(_double_factorial(2 * l - 1)) * math.pi ** 1.5 / float(2 ** l)
    K = c_l / np.power(alpha, float(l) + 1.5)
    s = float(c @ (K @ c))
    if not np.isfinite(s) or s <= 0.0:
        raise ValueError("non-positive contraction self-overlap; check coefficients")
    target = 1.0 if l <= 1 else 4.0 * math.pi / float(2 * l + 1)
    return c * math.sqrt(target / s)


# ---------------------------------------------------------------------------
# Parsing and packing
# ---------------------------------------------------------------------------


def _parse_entry(entry: Any) -> list[tuple[int, np.ndarray, np.ndarray]]:
    if not isinstance(entry, (list, tuple)) or len(entry) < 2:
        raise ValueError(f"invalid basis shell entry: {entry!r}")
    ls: list[int] = []
    start = None
    for i, item in enumerate(entry):
        if isinstance(item, (int, np.integer)):
            ls.append(int(item))
            continue
        start = i
        break
    if start is None or not ls:
        raise ValueError(f"invalid basis shell entry header: {entry!r}")
    rows = [list(map(float, line)) for line in entry[start:]]
    exps = np.asarray([r[0] for r in rows], dtype=np.float64)
    coef = np.asarray([r[1:] for r in rows], dtype=np.float64)
    if coef.ndim != 2 or coef.shape[0] != exps.size:
        raise ValueError("unexpected coefficient matrix shape")
    nL = len(ls)
    if coef.shape[1] % nL != 0:
        raise ValueError("coefficient column count not divisible by the number of l values")
    nctr = coef.shape[1] // nL
    return [(l, exps, coef[:, i * nctr:(i + 1) * nctr]) for i, l in enumerate(ls)]


def parse_basis_dict(basis: dict, elements: list[str]) -> dict[str, list[tuple[int, np.ndarray, np.ndarray]]]:
    """Per-element (l, exps, coefs) shells, ordered by l (stable)."""
    if not isinstance(basis, dict):
        raise TypeError("basis must be a dict mapping element symbol -> shell list")
    out: dict[str, list] = {}
    for sym in elements:
        spec = basis.get(sym)
        if spec is None:
            raise KeyError(f"missing basis for element {sym!r}")
        shells = []
        for entry in spec:
            shells.extend(_parse_entry(entry))
        shells = [s for _i, s in sorted(enumerate(shells), key=lambda t: (t[1][0], t[0]))]
        out[sym] = shells
    return out


@dataclass(frozen=True)
class BasisCartSoA:
    """Packed Cartesian basis (cueri/basis_cart.py:11-77 field layout)."""

    shell_cxyz: np.ndarray  # (nshell, 3) f64, Bohr
    shell_prim_start: np.ndarray  # (nshell,) i32
    shell_nprim: np.ndarray  # (nshell,) i32
    shell_l: np.ndarray  # (nshell,) i32
    shell_ao_start: np.ndarray  # (nshell,) i32
    prim_exp: np.ndarray  # (nprim,) f64
    prim_coef: np.ndarray  # (nprim,) f64, normalisation folded in
    shell_atom: np.ndarray | None = None  # (nshell,) i32, owning atom

    def __post_init__(self) -> None:
        n = int(self.shell_cxyz.shape[0])
        if self.shell_cxyz.dtype != np.float64 or self.shell_cxyz.shape != (n, 3):
            raise ValueError("shell_cxyz must be float64 with shape (nShell, 3)")
        for name in ("shell_prim_start", "shell_nprim", "shell_l", "shell_ao_start"):
            arr = getattr(self, name)
            if arr.dtype != np.int32 or arr.shape != (n,):
                raise ValueError(f"{name} must be int32 with shape (nShell,)")
        if self.prim_exp.dtype != np.float64 or self.prim_exp.shape != self.prim_coef.shape:
            raise ValueError("prim_exp/prim_coef must be float64 arrays of identical shape")

    @property
    def nshell(self) -> int:
        return int(self.shell_l.shape[0])

    @property
    def nao(self) -> int:
        if self.nshell == 0:
            return 0
        last = self.nshell - 1
        return int(self.shell_ao_start[last]) + ncart(int(self.shell_l[last]))


def pack_cart_basis(atoms_bohr, basis_shells: dict) -> BasisCartSoA:
    """Pack per-element shells into a :class:`BasisCartSoA`, one shell per
    contraction column (frontend/basis_packer.py:183-251)."""
    cxyz, pstart, nprim_l, ls, aostart, pexp, pcoef, satom = [], [], [], [], [], [], [], []
    ao = 0
    for iatm, (sym, xyz) in enumerate(atoms_bohr):
        xyz = np.asarray(xyz, dtype=np.float64).reshape(3)
        shells = basis_shells.get(str(sym).strip())
        if shells is None:
            raise KeyError(f"missing basis shells for element {sym!r}")
        for l, exps, coefs in shells:
            norm = primitive_norm(l, exps)
            for ctr in range(coefs.shape[1]):
                cxyz.append(xyz)
                pstart.append(len(pexp))
                nprim_l.append(int(exps.size))
                ls.append(int(l))
                aostart.append(ao)
                satom.append(iatm)
                prim = normalize_contraction(l, exps, coefs[:, ctr] * norm)
                pexp.extend(exps.tolist())
                pcoef.extend(prim.tolist())
                ao += ncart(l)
    return BasisCartSoA(
        shell_cxyz=np.asarray(cxyz, dtype=np.float64).reshape(-1, 3),
        shell_prim_start=np.asarray(pstart, dtype=np.int32),
        shell_nprim=np.asarray(nprim_l, dtype=np.int32),
        shell_l=np.asarray(ls, dtype=np.int32),
        shell_ao_start=np.asarray(aostart, dtype=np.int32),
        prim_exp=np.asarray(pexp, dtype=np.float64),
        prim_coef=np.asarray(pcoef, dtype=np.float64),
        shell_atom=np.asarray(satom, dtype=np.int32),
    )


def build_ao_basis_cart(mol, basis: Any | None = None) -> tuple[BasisCartSoA, str]:
    """(ao_basis, basis_name) for a molecule (frontend/one_electron.py:20-41).

    ``basis`` may be a vendored name ("cc-pVDZ", "cc-pVTZ") or an explicit dict.
    """
    spec = mol.basis if basis is None else basis
    elements = sorted({str(s) for s, _ in mol.atoms_bohr})
    if isinstance(spec, str):
        shells = parse_basis_dict(named_basis(spec), elements)
        name = spec
    elif isinstance(spec, dict):
        shells = parse_basis_dict(spec, elements)
        name = "<explicit>"
    else:
        raise TypeError("basis must be a string name or an explicit per-element basis dict")
    return pack_cart_basis(list(mol.atoms_bohr), shells), name


__all__ = [
    "ANGSTROM_TO_BOHR",
    "BasisCartSoA",
    "CC_PVDZ",
    "CC_PVTZ",
    "atomic_number",
    "build_ao_basis_cart",
    "cartesian_components",
    "named_basis",
    "ncart",
    "normalize_contraction",
    "pack_cart_basis",
    "parse_basis_dict",
    "primitive_norm",
]
