"""Molecules and the seeded synthetic water clusters used by tests and benchmarks.

``Molecule`` keeps the reference's container shape (frontend/molecule.py:51-165):
atoms in Bohr, charge, basis, cart flag, nuclear repulsion energy and its
gradient.

Water clusters: n molecules on the first n sites, in lexicographic order, of a
k^3 cubic grid (k = ceil(n^(1/3)), 3.0 Angstrom spacing). Each molecule is the
monomer O(0,0,0), H(0,+-0.7572,0.5865) Angstrom rotated about its O atom by a
Haar-random rotation and then shifted by a uniform jitter in [-0.3, 0.3] Angstrom
on each axis, all drawn from ``numpy.random.default_rng(seed)`` in molecule order
(rotation first, then jitter).
"""

from __future__ import annotations

import math
from dataclasses import dataclass, field
from typing import Any

import numpy as np

from .basis import ANGSTROM_TO_BOHR, atomic_number

WATER_MONOMER_ANG = (
    ("O", (0.0, 0.0, 0.0)),
    ("H", (0.0, 0.7572, 0.5865)),
    ("H", (0.0, -0.7572, 0.5865)),
)


@dataclass(frozen=True)
class Molecule:
    atoms_bohr: tuple[tuple[str, np.ndarray], ...]
    charge: int = 0
    spin: int = 0
    basis: Any = None
    cart: bool = True
    results: dict = field(default_factory=dict, compare=False, repr=False)

    @classmethod
    def from_atoms(cls, atoms, *, unit: str = "Bohr", charge: int = 0, spin: int = 0,
                   basis: Any = None, cart: bool = True) -> "Molecule":
        u = str(unit).strip().lower()
        if u in ("bohr", "a0", "au"):
            scale = 1.0
        elif u in ("angstrom", "ang", "a"):
            scale = ANGSTROM_TO_BOHR
        else:
            raise ValueError(f"unknown unit {unit!r}")
        parsed = []
        if isinstance(atoms, str):
            for frag in atoms.replace("\n", ";").split(";"):
                tok = frag.split()
                if not tok:
                    continue
                if len(tok) != 4:
                    raise ValueError(f"invalid atom fragment {frag!r}")
                parsed.append((tok[0], np.array([float(t) for t in tok[1:]], dtype=np.float64)))
        else:
            for sym, xyz in atoms:
                parsed.append((str(sym), np.asarray(xyz, dtype=np.float64).reshape(3).copy()))
        if not parsed:
            raise ValueError("no atoms parsed")
        return cls(tuple((s, x * scale) for s, x in parsed), charge=int(charge), spin=int(spin),
                   basis=basis, cart=bool(cart))

    @property
    def natm(self) -> int:
        return len(self.atoms_bohr)

    @property
    def elements(self) -> list[str]:
        return [s for s, _ in self.atoms_bohr]

    @property
    def coords_bohr(self) -> np.ndarray:
        return np.asarray([x for _, x in self.atoms_bohr], dtype=np.float64).reshape(-1, 3)

    @property
    def charges(self) -> np.ndarray:
        return np.asarray([atomic_number(s) for s, _ in self.atoms_bohr], dtype=np.float64)

    @property
    def nelectron(self) -> int:
        return int(sum(atomic_number(s) for s, _ in self.atoms_bohr) - int(self.charge))

    def with_coords(self, coords_bohr: np.ndarray) -> "Molecule":
        c = np.asarray(coords_bohr, dtype=np.float64).reshape(self.natm, 3)
        atoms = tuple((s, c[i].copy()) for i, (s, _) in enumerate(self.atoms_bohr))
        return Molecule(atoms, charge=self.charge, spin=self.spin, basis=self.basis, cart=self.cart)

    def energy_nuc(self) -> float:
        """Nuclear repulsion energy (frontend/molecule.py:131-146)."""
        R, Z = self.coords_bohr, self.charges
        e = 0.0
        for i in range(self.natm):
            for j in range(i + 1, self.natm):
                r = float(np.linalg.norm(R[i] - R[j]))
                if r == 0.0:
                    raise ValueError("coincident nuclei")
                e += Z[i] * Z[j] / r
        return float(e)

    def energy_nuc_grad(self) -> np.ndarray:
        """d E_nuc / d R, shape (natm, 3) (frontend/molecule.py:148-165)."""
        R, Z = self.coords_bohr, self.charges
        g = np.zeros((self.natm, 3))
        for i in range(self.natm):
            for j in range(i + 1, self.natm):
                d = R[i] - R[j]
                r = float(np.linalg.norm(d))
                f = -Z[i] * Z[j] / r ** 3
                g[i] += f * d
                g[j] -= f * d
        return g


def haar_rotation(rng: np.random.Generator) -> np.ndarray:
    """Uniform (Haar) random rotation: QR of a Gaussian matrix with sign fix."""
    A = rng.standard_normal((3, 3))
    Qm, Rm = np.linalg.qr(A)
    Qm = Qm * np.sign(np.diag(Rm))[None, :]
    if np.linalg.det(Qm) < 0:
        Qm[:, 0] = -Qm[:, 0]
    return Qm


def water_cluster_atoms_ang(n: int, seed: int, spacing: float = 3.0, jitter: float = 0.3,
                            grid: int | None = None):
    """Atom list (symbol, xyz in Angstrom) of the seeded n-water cluster."""
    if n < 1:
        raise ValueError("n must be >= 1")
    k = int(grid) if grid is not None else int(math.ceil(round(n ** (1.0 / 3.0), 12)))
    while k ** 3 < n:
        k += 1
    sites = [(i, j, l) for i in range(k) for j in range(k) for l in range(k)][:n]
    rng = np.random.default_rng(int(seed))
    mono = np.array([x for _, x in WATER_MONOMER_ANG], dtype=np.float64)
    atoms = []
    for site in sites:
        rot = haar_rotation(rng)
        shift = rng.uniform(-jitter, jitter, size=3)
        origin = spacing * np.asarray(site, dtype=np.float64) + shift
        xyz = mono @ rot.T + origin
        for (sym, _), pos in zip(WATER_MONOMER_ANG, xyz):
            atoms.append((sym, pos))
    return atoms


def water_cluster(n: int, seed: int, basis: Any = "cc-pVDZ", grid: int | None = None) -> Molecule:
    return Molecule.from_atoms(water_cluster_atoms_ang(n, seed, grid=grid), unit="Angstrom", basis=basis)


def water_monomer(basis: Any = "cc-pVDZ") -> Molecule:
    return Molecule.from_atoms(list(WATER_MONOMER_ANG), unit="Angstrom", basis=basis)


__all__ = ["Molecule", "haar_rotation", "water_cluster", "water_cluster_atoms_ang", "water_monomer"]
