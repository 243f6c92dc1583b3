"""The oracle's UHF/ROHF restatement (hf/direct_scf.py:405-, hf/df_scf.py:691-719),
pinned by what the mathematics fixes:

* a closed shell (nalpha = nbeta) gives the RHF energy of the oracle's RHF;
* one electron (H atom doublet): the self-Coulomb and self-exchange of the
  single occupied orbital cancel exactly, so UHF and ROHF energies are the
  lowest generalised eigenvalue of (h, S);
* ROHF is a constrained UHF, so E_ROHF >= E_UHF (OH radical doublet), and the
  two differ (the open shell polarises the UHF core).
"""

import numpy as np
import pytest
import scipy.linalg

from asuka_amd.basis import build_ao_basis_cart
from asuka_amd.molecule import Molecule, water_monomer


def _setup(oracle, mol):
    ao, _ = build_ao_basis_cart(mol)
    octx = oracle.make_context(ao)
    eri = oracle.dense_eri(octx, hp=False)
    S, T, V = oracle.int1e(ao, mol.coords_bohr, mol.charges)

    def jk(D):
        J, K = oracle.dense_JK(eri, D)
        return 0.5 * (J + J.T), 0.5 * (K + K.T)

    return S, T + V, jk


def test_closed_shell_open_shell_scf_is_rhf(oracle):
    mol = water_monomer("cc-pVDZ")
    S, h, jk = _setup(oracle, mol)

    def fock(D):
        J, K = jk(D)
        return h + J - 0.5 * K

    ref = oracle.rhf(S, h, fock, mol.nelectron, mol.energy_nuc(), conv_tol=1e-12, conv_tol_dm=1e-10)
    n = mol.nelectron // 2
    for rohf in (False, True):
        r = oracle.open_shell_scf(S, h, jk, n, n, mol.energy_nuc(), rohf=rohf)
        assert r["converged"]
        assert abs(r["e_tot"] - ref["e_tot"]) < 1e-9, (rohf, r["e_tot"], ref["e_tot"])


def test_one_electron_energy_is_lowest_eigenvalue(oracle):
    mol = Molecule.from_atoms([("H", (0.0, 0.0, 0.0))], spin=1, basis="cc-pVDZ")
    S, h, jk = _setup(oracle, mol)
    e0 = scipy.linalg.eigh(h, S, eigvals_only=True)[0]
    for rohf in (False, True):
        r = oracle.open_shell_scf(S, h, jk, 1, 0, 0.0, rohf=rohf)
        assert r["converged"]
        assert abs(r["e_tot"] - e0) < 1e-10, (rohf, r["e_tot"], e0)


@pytest.fixture(scope="module")
def oh_radical():
    return Molecule.from_atoms([("O", (0.0, 0.0, 0.0)), ("H", (0.0, 0.0, 0.97))], unit="Angstrom", spin=1,
                               basis="cc-pVDZ")


def test_rohf_above_uhf(oracle, oh_radical):
    mol = oh_radical
    S, h, jk = _setup(oracle, mol)
    na, nb = (mol.nelectron + 1) // 2, (mol.nelectron - 1) // 2
    u = oracle.open_shell_scf(S, h, jk, na, nb, mol.energy_nuc(), max_cycle=400)
    r = oracle.open_shell_scf(S, h, jk, na, nb, mol.energy_nuc(), rohf=True, max_cycle=400)
    assert u["converged"] and r["converged"]
    assert r["e_tot"] > u["e_tot"] + 1e-6
    assert r["e_tot"] - u["e_tot"] < 0.05  # spin contamination of a doublet radical is small
