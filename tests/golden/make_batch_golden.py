"""Write ``batch_golden.npz``: oracle J/K (binary128 roots) for the small
clusters of the batched workload.

The batch is 32 water clusters, cc-pVDZ cartesian, sizes
``default_rng(3).integers(1, 17, 32)``, cluster i placed with seed 1000 + i
and given D = 2 C C^T with C = QR(N(0,1) nao x 5n, seed 2000 + i). For every
cluster of at most 4 waters this stores J and K (symmetrised) at Schwarz 1e-12
and density screening 1e-10 (the reference defaults, asuka/hf/direct_jk.py:
1228-1246, :1594-1604, :1893-1909) as ``J_<i>``, ``K_<i>``.
Run: ``python tests/golden/make_batch_golden.py`` (minutes, all cores).
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

MAX_WATERS = 4


def batch_inputs(i: int, n: int):
    ao, _ = build_ao_basis_cart(water_cluster(int(n), 1000 + i, "cc-pVDZ"))
    C = np.random.default_rng(2000 + i).standard_normal((ao.nao, 5 * int(n)))
    C, _ = np.linalg.qr(C)
    return ao, 2.0 * C @ C.T


def main():
    sizes = np.random.default_rng(3).integers(1, 17, 32)
    out = {"sizes": sizes}
    for i, n in enumerate(sizes):
        if n > MAX_WATERS:
            continue
        ao, D = batch_inputs(i, n)
        ctx = O.make_context(ao)
        J, K = O.direct_jk(ctx, D, eps_schwarz=1e-12, hp=True)
        out[f"J_{i}"] = J
        out[f"K_{i}"] = K
        print(i, int(n), ao.nao, flush=True)
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "batch_golden.npz")
    np.savez_compressed(path, **out)
    print("wrote", path)


if __name__ == "__main__":
    main()
