"""The autotuner and tuning tables (asuka_jk_autotune, asuka_jk_tuning_get/set).

Launch shapes change only how work is cut and where it is summed, never
which quartets are evaluated, so every shape must reproduce the oracle's
golden J/K to the FP64 bar (1e-10 absolute): the tuned shapes, and two
forced extremes applied to every class pair (8 wavefronts per workgroup, 8
wave units per task, no LDS accumulators; 4 wavefronts, 64 units, LDS
accumulators where they fit).
"""

import os

import numpy as np
import pytest

from asuka_amd.basis import build_ao_basis_cart
from asuka_amd.molecule import water_cluster

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
CASES = {"h2o2_tz": (2, 2, "cc-pVTZ"), "h2o4_dz": (4, 4, "cc-pVDZ")}


def _ctx(mol, **kw):
    from asuka_amd.hf.direct_jk import make_direct_jk_context

    ao, _ = build_ao_basis_cart(mol)
    return ao, make_direct_jk_context(ao, atom_coords_bohr=mol.coords_bohr, atom_charges=mol.charges, **kw)


@pytest.fixture(scope="module")
def jk_golden():
    return dict(np.load(os.path.join(HERE, "golden", "jk_golden.npz"), allow_pickle=False))


def _check(ctx, D, golden, case):
    from asuka_amd.hf.direct_jk import direct_JK

    J, K = direct_JK(ctx, D)
    assert np.max(np.abs(J.cpu().numpy() - golden[f"{case}_J_scr"])) < 1e-10
    assert np.max(np.abs(K.cpu().numpy() - golden[f"{case}_K_scr"])) < 1e-10


@pytest.mark.parametrize("case", sorted(CASES))
def test_autotune_keeps_parity(gpu, jk_golden, case):
    n, seed, basis = CASES[case]
    mol = water_cluster(n, seed, basis)
    D = jk_golden[f"{case}_D"]
    _, ctx = _ctx(mol, eps_schwarz=1e-12)
    table = ctx.autotune(D, reps=1)
    assert len(table) > 0 and all(t["tuned_ms"] > 0 for t in table)
    assert {t["waves_per_group"] for t in table} <= {4, 8}
    assert all(t["units_per_task"] in (16, 32, 64, 128) for t in table)  # a candidate or the default cut
    _check(ctx, D, jk_golden, case)
    # the table carries over to a fresh context of the same molecule
    _, ctx2 = _ctx(mol, eps_schwarz=1e-12)
    ctx2.set_tuning(table)
    t2 = ctx2.tuning()
    assert [(t["class"], t["waves_per_group"], t["units_per_task"]) for t in t2] == \
        [(t["class"], t["waves_per_group"], t["units_per_task"]) for t in table]
    _check(ctx2, D, jk_golden, case)


@pytest.mark.parametrize("case", sorted(CASES))
def test_forced_launch_shapes(gpu, jk_golden, case):
    n, seed, basis = CASES[case]
    mol = water_cluster(n, seed, basis)
    D = jk_golden[f"{case}_D"]
    _, ctx = _ctx(mol, eps_schwarz=1e-12)
    classes = [t["class"] for t in ctx.tuning()]
    for wpg, tu, acc in ((8, 8, False), (4, 64, True), (8, 32, True), (4, 16, False)):
        ctx.set_tuning([{"class": c, "waves_per_group": wpg, "units_per_task": tu, "lds_accumulators": acc,
                         "tuned_ms": 0.0} for c in classes])
        _check(ctx, D, jk_golden, case)


def test_tuning_errors(gpu):
    mol = water_cluster(2, 3, "cc-pVDZ")
    _, ctx = _ctx(mol)
    c = ctx.tuning()[0]["class"]
    with pytest.raises(ValueError, match="must be"):
        ctx.set_tuning([{"class": c, "waves_per_group": 6, "units_per_task": 16, "lds_accumulators": True,
                         "tuned_ms": 0.0}])
    with pytest.raises(ValueError, match="must be"):
        ctx.set_tuning([{"class": c, "waves_per_group": 4, "units_per_task": 0, "lds_accumulators": True,
                         "tuned_ms": 0.0}])
    with pytest.raises(ValueError, match="must be"):
        ctx.autotune(np.zeros((3, 3)))
