"""Write ``jk_golden.npz``: oracle J/K (binary128 roots) for seeded water clusters.

Cases (seeded inputs, all recorded in the file):
  h2o2_tz : (H2O)2, seed 2, cc-pVTZ cartesian (130 AOs)
  h2o4_dz : (H2O)4, seed 4, cc-pVDZ cartesian (100 AOs)
For each: D = C C^T with C = QR(N(0,1) nao x 10, seed 11), and J, K
(symmetrised) for
  *_scr  : Schwarz 1e-12 and density screening at 1e-10 (the reference defaults,
           asuka/hf/direct_jk.py:1228-1246, :1594-1604, :1893-1909);
  *_full : no screening.
Run: ``python tests/golden/make_jk_golden.py`` (several minutes, all cores).
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

CASES = {"h2o2_tz": (2, 2, "cc-pVTZ"), "h2o4_dz": (4, 4, "cc-pVDZ")}


def density(nao: int, nocc: int = 10, seed: int = 11) -> np.ndarray:
    C = np.random.default_rng(seed).standard_normal((nao, nocc))
    C, _ = np.linalg.qr(C)
    return C @ C.T


def main():
    out = {}
    for name, (n, seed, basis) in CASES.items():
        mol = water_cluster(n, seed, basis)
        ao, _ = build_ao_basis_cart(mol)
        ctx = O.make_context(ao)
        D = density(ao.nao)
        Q = O.schwarz(ctx, hp=True)
        J, K = O.direct_jk(ctx, D, eps_schwarz=1e-12, Q=Q, hp=True)
        out[f"{name}_D"] = D
        out[f"{name}_Q"] = Q
        out[f"{name}_J_scr"] = J
        out[f"{name}_K_scr"] = K
        J, K = O.direct_jk(ctx, D, eps_schwarz=0.0, eps_density=0.0, hp=True)
        out[f"{name}_J_full"] = J
        out[f"{name}_K_full"] = K
        print(name, ao.nao, flush=True)
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "jk_golden.npz")
    np.savez_compressed(path, **out)
    print("wrote", path)


if __name__ == "__main__":
    main()
