"""The SCF iteration captured in one HIP graph (rhf_direct, use_graph=True).

* asuka_scf_diis_solve: its weights, applied to the stored Fock matrices, give
  the same extrapolated Fock matrix as the host DIIS (numpy solve of the
  reference's bordered system), for 1..8 stored vectors and after the ring
  wraps;
* graph replays equal the same cycle run eagerly, cycle by cycle;
* the graph SCF (density by purification, device DIIS) follows the eager
  eigh-based SCF cycle by cycle and converges to the same energy, density and
  orbital energies;
* the converged energy equals the oracle's RHF energy on dense ERIs.
"""

import ctypes

import numpy as np
import pytest

from asuka_amd.basis import build_ao_basis_cart
from asuka_amd.molecule import water_cluster, water_monomer

pytestmark = pytest.mark.gpu


def test_device_diis_matches_host_diis(gpu):
    from asuka_amd import _lib
    from asuka_amd.hf.direct_scf import _DIIS

    torch = gpu
    n = 12
    rng = np.random.default_rng(5)
    for nslot in (8, 5):
        host = _DIIS(torch, nslot)
        Fs = torch.zeros((8, n, n), dtype=torch.float64, device="cuda")
        Es = torch.zeros((8, n * n), dtype=torch.float64, device="cuda")
        G = torch.zeros((8, 8), dtype=torch.float64, device="cuda")
        count = torch.zeros(1, dtype=torch.int64, device="cuda")
        coeff = torch.zeros(8, dtype=torch.float64, device="cuda")
        status = torch.zeros(1, dtype=torch.int32, device="cuda")
        for k in range(11):
            F = torch.tensor(rng.standard_normal((n, n)), device="cuda")
            e = torch.tensor(rng.standard_normal((n, n)) * 0.5 ** k, device="cuda")
            host.push(F, e)
            want = host.extrapolate()
            slot = torch.remainder(count, nslot)
            Fs.index_copy_(0, slot, F[None])
            Es.index_copy_(0, slot, e.reshape(1, -1))
            g = Es @ e.reshape(-1)
            G.index_copy_(0, slot, g[None])
            G.index_copy_(1, slot, g[:, None])
            count.add_(1)
            _lib.check(_lib.lib().asuka_scf_diis_solve(_lib.ptr(G), _lib.ptr(count), nslot, _lib.ptr(coeff),
                                                        _lib.ptr(status), ctypes.c_void_p(_lib.stream_handle())))
            got = torch.tensordot(coeff, Fs, dims=([0], [0]))
            assert int(status.item()) == (1 if k >= 1 else 0)
            assert abs(float(coeff.sum()) - 1.0) < 1e-12
            assert float((got - want).abs().max()) < 1e-9 * max(1.0, float(want.abs().max())), (nslot, k)


def _scf(mol, **kw):
    from asuka_amd.frontend import run_hf_df

    prof = {}
    r = run_hf_df(mol, two_e_backend="direct", eps_schwarz=1e-12, profile=prof, **kw)
    return r, prof["scf"]


def test_graph_scf_follows_eager_scf(gpu, monkeypatch):
    from asuka_amd.hf import direct_scf

    mol = water_cluster(4, 4, "cc-pVDZ")
    print("\n[scf-graph] (H2O)4 cc-pVDZ eager", flush=True)
    re, pe = _scf(mol, use_graph=False, conv_tol=1e-10)
    print("[scf-graph] graph", flush=True)
    rg, pg = _scf(mol, use_graph=True, conv_tol=1e-10)
    assert pe["graph"] is False and pg["graph"] is True
    assert re.scf.converged and rg.scf.converged
    assert abs(rg.scf.e_tot - re.scf.e_tot) < 1e-9
    assert abs(rg.scf.niter - re.scf.niter) <= 1
    assert pg["purify_err_max"] < 1e-10
    m = min(len(pe["energies"]), len(pg["energies"]))
    assert m >= 5
    assert np.abs(np.array(pe["energies"][:m]) - np.array(pg["energies"][:m])).max() < 1e-8
    assert float((rg.scf.dm - re.scf.dm).abs().max()) < 1e-6
    nocc = mol.nelectron // 2
    assert float((rg.scf.mo_energy[:nocc + 5] - re.scf.mo_energy[:nocc + 5]).abs().max()) < 1e-6
    print("[scf-graph] the same cycle run eagerly", flush=True)
    monkeypatch.setattr(direct_scf, "_GRAPH_REPLAY", False)
    rx, px = _scf(mol, use_graph=True, conv_tol=1e-10)
    assert rx.scf.niter == rg.scf.niter
    assert np.abs(np.array(px["energies"]) - np.array(pg["energies"])).max() < 1e-11


def test_graph_scf_fixed_iterations_and_oracle(gpu, oracle):
    """15 cycles with the convergence test off run to the end; the converged
    graph SCF of the water monomer equals the oracle's dense RHF energy."""
    mol = water_cluster(2, 3, "cc-pVDZ")
    r, p = _scf(mol, use_graph=True, conv_tol=0.0, conv_tol_dm=0.0, max_cycle=15, check_every=4)
    assert not r.scf.converged and r.scf.niter == 15 and len(p["energies"]) == 15
    mol = water_monomer("cc-pVDZ")
    r, _ = _scf(mol, use_graph=True)
    assert r.scf.converged
    ao, _ = build_ao_basis_cart(mol)
    octx = oracle.make_context(ao)
    eri = oracle.dense_eri(octx, hp=True)
    S, T, V = oracle.int1e(ao, mol.coords_bohr, mol.charges)
    h = T + V

    def fock(D):
        J, K = oracle.dense_JK(eri, D)
        return 0.5 * ((h + J - 0.5 * K) + (h + J - 0.5 * K).T)

    ref = oracle.rhf(S, h, fock, mol.nelectron, mol.energy_nuc())
    assert abs(r.scf.e_tot - ref["e_tot"]) < 1e-9
