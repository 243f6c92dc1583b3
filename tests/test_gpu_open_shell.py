"""UHF and ROHF through run_hf_df(method=..., two_e_backend="direct") over the
two-density build, against the oracle's open-shell SCF on dense FP64 ERIs
(1e-9 Eh at convergence), plus the closed-shell and one-electron limits."""

import numpy as np
import pytest
import scipy.linalg

from asuka_amd.basis import build_ao_basis_cart
from asuka_amd.molecule import Molecule, water_monomer

pytestmark = pytest.mark.gpu

KW = dict(two_e_backend="direct", eps_schwarz=1e-14, conv_tol=1e-12, conv_tol_dm=1e-10, max_cycle=400)


def _oh():
    return Molecule.from_atoms([("O", (0.0, 0.0, 0.0)), ("H", (0.0, 0.0, 0.97))], unit="Angstrom", spin=1,
                               basis="cc-pVDZ")


def _oracle_setup(oracle, mol):
    ao, _ = build_ao_basis_cart(mol)
    eri = oracle.dense_eri(oracle.make_context(ao), hp=True)
    S, T, V = oracle.int1e(ao, mol.coords_bohr, mol.charges)

    def jk(D):
        J, K = oracle.dense_JK(eri, D)
        return 0.5 * (J + J.T), 0.5 * (K + K.T)

    return S, T + V, jk


@pytest.mark.parametrize("method", ["uhf", "rohf"])
def test_oh_radical_against_oracle(gpu, oracle, method):
    from asuka_amd.frontend import run_hf_df

    mol = _oh()
    res = run_hf_df(mol, method=method, **KW)
    assert res.scf.converged and res.scf.method == method.upper()
    S, h, jk = _oracle_setup(oracle, mol)
    na, nb = (mol.nelectron + 1) // 2, (mol.nelectron - 1) // 2
    ref = oracle.open_shell_scf(S, h, jk, na, nb, mol.energy_nuc(), rohf=(method == "rohf"), max_cycle=400)
    assert ref["converged"]
    assert abs(res.scf.e_tot - ref["e_tot"]) < 1e-9, (res.scf.e_tot, ref["e_tot"])
    Da, Db = res.scf.dm
    assert abs(float((Da.cpu().numpy() * S).sum()) - na) < 1e-8
    assert abs(float((Db.cpu().numpy() * S).sum()) - nb) < 1e-8


def test_h_atom_and_closed_shell_limits(gpu):
    from asuka_amd.frontend import build_int1e, run_hf_df

    h_atom = Molecule.from_atoms([("H", (0.0, 0.0, 0.0))], spin=1, basis="cc-pVDZ")
    for method in ("uhf", "rohf"):
        res = run_hf_df(h_atom, method=method, **KW)
        ints = res.int1e
        e0 = scipy.linalg.eigh(ints.hcore.cpu().numpy(), ints.S.cpu().numpy(), eigvals_only=True)[0]
        assert abs(res.scf.e_tot - e0) < 1e-10, (method, res.scf.e_tot, e0)
    mol = water_monomer("cc-pVDZ")
    e_rhf = run_hf_df(mol, **KW).scf.e_tot
    for method in ("uhf", "rohf"):
        assert abs(run_hf_df(mol, method=method, **KW).scf.e_tot - e_rhf) < 1e-9, method
    with pytest.raises(NotImplementedError):
        run_hf_df(mol, method="uks", two_e_backend="direct")
    with pytest.raises(ValueError, match="nalpha"):
        run_hf_df(Molecule.from_atoms([("H", (0.0, 0.0, 0.0))], spin=0, basis="cc-pVDZ"), method="uhf",
                  two_e_backend="direct")
