"""Write ``grad_golden.npz``: the oracle's two-electron gradient of (H2O)4
cc-pVDZ cartesian (seed 4).

D = 2 C C^T with C = QR(N(0,1) 100 x 20, seed 17), adjoints WJ = D/2,
WK = -D/4 (the RHF two-electron energy), gradient = d/dR [Tr(WJ J(D)) +
Tr(WK K(D))] at fixed D over all ordered shell-pair pairs, the reference
driver of asuka/mcscf/nac/_dense.py:1513-1839 with the derivative kernel of
asuka/cueri/_eri_rys_cpu.pyx:4316-4656 (oracle.grad2e in the reference's double-precision
arithmetic; its root error, below 1e-10 relative, is far inside the test's 1e-9 bar).
Stored: ``D``, ``grad``.
Run: ``python tests/golden/make_grad_golden.py`` (minutes, all cores).
"""

from __future__ import annotations

import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

from asuka_amd.basis import build_ao_basis_cart  # noqa: E402
from asuka_amd.molecule import water_cluster  # noqa: E402
from oracle import oracle as O  # noqa: E402


def main():
    mol = water_cluster(4, 4, "cc-pVDZ")
    ao, _ = build_ao_basis_cart(mol)
    C = np.random.default_rng(17).standard_normal((ao.nao, 20))
    C, _ = np.linalg.qr(C)
    D = 2.0 * C @ C.T
    octx = O.make_context(ao)
    g = O.grad2e(octx, D, 0.5 * D, -0.25 * D, ao.shell_atom, mol.natm, hp=False, threads=0)
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "grad_golden.npz")
    np.savez_compressed(path, D=D, grad=g)
    print("wrote", path, "max |g|", float(np.abs(g).max()), "translational sum", float(np.abs(g.sum(0)).max()))


if __name__ == "__main__":
    main()
