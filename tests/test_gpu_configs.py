"""GPU tests at the benchmark workloads' full sizes.

* (H2O)64 cc-pVDZ (1600 AOs), the Fock-build workload: the Schwarz-screened
  and evaluated quartet sets equal the oracle's per class pair (integer
  equality); exact symmetry; linearity in D with density screening off;
  F = h + J - K/2; the mixed build within 1e-6 Eh of the FP64 build.
* (H2O)8 cc-pVTZ (520 AOs), the J/K workload: the mixed leg against the FP64
  build at Schwarz 1e-12, and the Schwarz 1e-8 leg (its quartet set against
  the oracle's, mixed against FP64, and its J/K near the 1e-12 build).
* The batch of 32 clusters of 1-16 waters (sizes default_rng(3).integers(1,
  17, 32), placement seeds 1000 + i): the batch context equals per-molecule
  contexts, clusters of up to 4 waters equal the oracle's J/K, and graph
  replays equal the eager build.
* (H2O)16 cc-pVDZ at the converged RHF density: the autograd Function's dL/dR
  equals asuka_jk_grad, the full gradient sums to zero over atoms, and central
  differences (1e-4 Bohr) on 2 of the 6 seeded coordinates agree to 1e-6
  Eh/Bohr.

Sizes follow the workloads exactly; every test prints its case before the
first GPU call.
"""

import os

import numpy as np
import pytest

from asuka_amd.basis import build_ao_basis_cart
from asuka_amd.molecule import water_cluster

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))


def _orth_density(nao, nocc, seed, scale=2.0):
    C = np.random.default_rng(seed).standard_normal((nao, nocc))
    C, _ = np.linalg.qr(C)
    return scale * C @ C.T


def _ctx(mol, **kw):
    from asuka_amd.hf.direct_jk import make_direct_jk_context

    ao, _ = build_ao_basis_cart(mol)
    return ao, make_direct_jk_context(ao, atom_coords_bohr=mol.coords_bohr, atom_charges=mol.charges, **kw)


def test_fock_workload_h2o64_dz(gpu, oracle):
    from asuka_amd.frontend import build_int1e
    from asuka_amd.hf.direct_jk import direct_fock_rhf, direct_JK
    from test_gpu_screening import check_counts

    torch = gpu
    print("\n[configs] (H2O)64 cc-pVDZ: context", flush=True)
    mol = water_cluster(64, 5, "cc-pVDZ")
    ao, ctx = _ctx(mol, eps_schwarz=1e-12)
    assert ao.nao == 1600
    D1 = _orth_density(ao.nao, 320, 1005)
    print("[configs] (H2O)64 cc-pVDZ: quartet counts", flush=True)
    check_counts(ctx, oracle, oracle.make_context(ao), D1)
    h = build_int1e(ctx).hcore
    print("[configs] (H2O)64 cc-pVDZ: Fock and J/K builds", flush=True)
    J, K = direct_JK(ctx, D1)
    assert float((J - J.T).abs().max()) == 0.0 and float((K - K.T).abs().max()) == 0.0
    F = direct_fock_rhf(ctx, D1, h)
    ref = h + J - 0.5 * K
    assert float((F - ref).abs().max()) < 1e-10 * max(1.0, float(ref.abs().max()))
    Fm = direct_fock_rhf(ctx, D1, h, mixed_precision=True)
    D1t = torch.as_tensor(D1, device="cuda")
    dE = float(0.5 * (D1t * (Fm - F)).sum())
    assert abs(dE) < 1e-6, dE
    print("[configs] (H2O)64 cc-pVDZ: linearity (density screening off)", flush=True)
    ao, ctx0 = _ctx(mol, eps_schwarz=1e-12, eps_density=0.0)
    D2 = _orth_density(ao.nao, 320, 1006)
    J1, K1 = direct_JK(ctx0, D1)
    J2, K2 = direct_JK(ctx0, D2)
    J12, K12 = direct_JK(ctx0, D1 + 0.5 * D2)
    tol = 1e-12 * max(float(J1.abs().max()), float(K1.abs().max()), 1.0) * 100
    assert float((J12 - J1 - 0.5 * J2).abs().max()) < tol
    assert float((K12 - K1 - 0.5 * K2).abs().max()) < tol


def test_jk_workload_h2o8_tz_mixed_and_eps8(gpu, oracle):
    from asuka_amd.hf.direct_jk import direct_JK
    from test_gpu_screening import check_counts

    torch = gpu
    print("\n[configs] (H2O)8 cc-pVTZ: mixed vs FP64 at Schwarz 1e-12", flush=True)
    mol = water_cluster(8, 2, "cc-pVTZ")
    ao, ctx = _ctx(mol, eps_schwarz=1e-12)
    assert ao.nao == 520
    D = _orth_density(ao.nao, 40, 1002, scale=1.0)
    Dt = torch.as_tensor(D, device="cuda")
    J, K = direct_JK(ctx, D)
    Jm, Km = direct_JK(ctx, D, mixed_precision=True)
    assert float((Jm - J).abs().max()) < 1e-8 and float((Km - K).abs().max()) < 1e-8
    dE = float(0.5 * (Dt * (Jm - J)).sum() - 0.25 * (Dt * (Km - K)).sum())
    assert abs(dE) < 1e-8
    print("[configs] (H2O)8 cc-pVTZ: Schwarz 1e-8", flush=True)
    ao, ctx8 = _ctx(mol, eps_schwarz=1e-8)
    octx = oracle.make_context(ao)
    check_counts(ctx8, oracle, octx, D)
    J8, K8 = direct_JK(ctx8, D)
    J8m, K8m = direct_JK(ctx8, D, mixed_precision=True)
    assert float((J8m - J8).abs().max()) < 1e-8 and float((K8m - K8).abs().max()) < 1e-8
    # the 1e-8 build leaves out quartets with Q_ab Q_cd < 1e-8 (or, with the
    # density, below eps_density = 1e-8): fewer quartets, J/K close to the
    # 1e-12 build
    assert ctx8.ntasks < ctx.ntasks
    dJ = float((J8 - J).abs().max())
    dK = float((K8 - K).abs().max())
    print(f"[configs] Schwarz 1e-8 vs 1e-12: max |dJ| {dJ:.2e}, max |dK| {dK:.2e}", flush=True)
    assert 0.0 < dJ < 1e-4 and 0.0 < dK < 1e-4


def test_batched_workload_32_clusters(gpu):
    from asuka_amd.hf.direct_jk import direct_JK_batched, make_direct_jk_batch_context, make_direct_jk_context

    torch = gpu
    sizes = np.random.default_rng(3).integers(1, 17, 32)
    print(f"\n[configs] batch of 32 clusters, sizes {sizes.tolist()}", flush=True)
    aos, Ds = [], []
    for i, n in enumerate(sizes):
        ao, _ = build_ao_basis_cart(water_cluster(int(n), 1000 + i, "cc-pVDZ"))
        aos.append(ao)
        Ds.append(_orth_density(ao.nao, 5 * int(n), 2000 + i))
    bc = make_direct_jk_batch_context(aos, eps_schwarz=1e-12)
    assert bc.naos == [a.nao for a in aos]
    ctxs = [make_direct_jk_context(a, eps_schwarz=1e-12) for a in aos]
    ref = direct_JK_batched(ctxs, Ds)
    out = direct_JK_batched(bc, Ds)
    torch.cuda.synchronize()
    for (J, K), (J0, K0) in zip(out, ref):
        assert float((J - J0).abs().max()) < 1e-11 and float((K - K0).abs().max()) < 1e-11
    print("[configs] batch: clusters of <= 4 waters against the oracle's golden J/K", flush=True)
    gold = dict(np.load(os.path.join(HERE, "golden", "batch_golden.npz"), allow_pickle=False))
    assert np.array_equal(gold["sizes"], sizes)
    nchk = 0
    for i, n in enumerate(sizes):
        if f"J_{i}" not in gold:
            assert n > 4
            continue
        J, K = out[i]
        assert np.abs(J.cpu().numpy() - gold[f"J_{i}"]).max() < 1e-10, (i, int(n))
        assert np.abs(K.cpu().numpy() - gold[f"K_{i}"]).max() < 1e-10, (i, int(n))
        nchk += 1
    assert nchk == int((sizes <= 4).sum())
    print("[configs] batch: graph replays", flush=True)
    Ds = [torch.as_tensor(D, device="cuda") for D in Ds]  # no host copies inside the capture
    bufs = [(torch.empty_like(J), torch.empty_like(K)) for J, K in out]
    side = torch.cuda.Stream()
    side.wait_stream(torch.cuda.current_stream())
    with torch.cuda.stream(side):
        direct_JK_batched(bc, Ds, out=bufs)
    torch.cuda.current_stream().wait_stream(side)
    torch.cuda.synchronize()
    graph = torch.cuda.CUDAGraph()
    with torch.cuda.graph(graph):
        direct_JK_batched(bc, Ds, out=bufs)
    for rep in range(3):
        for J, K in bufs:
            J.fill_(7.0)
            K.fill_(-7.0)
        graph.replay()
        torch.cuda.synchronize()
        for (J, K), (J0, K0) in zip(bufs, out):
            assert float((J - J0).abs().max()) < 1e-11 and float((K - K0).abs().max()) < 1e-11, rep
        direct_JK_batched(bc, Ds)  # eager work between replays
        torch.cuda.synchronize()


def test_gradient_workload_h2o16_dz(gpu):
    from asuka_amd.frontend import run_hf_df
    from asuka_amd.hf.grad import direct_jk_autograd, direct_jk_grad, rhf_nuc_grad

    torch = gpu
    print("\n[configs] (H2O)16 cc-pVDZ: SCF", flush=True)
    mol = water_cluster(16, 4, "cc-pVDZ")
    res = run_hf_df(mol, two_e_backend="direct", eps_schwarz=1e-12, conv_tol=1e-10)
    assert res.scf.converged
    ctx = res.direct_jk_ctx
    print("[configs] (H2O)16 cc-pVDZ: autograd vs asuka_jk_grad", flush=True)
    D = res.scf.dm.detach().clone().requires_grad_(True)
    R = torch.tensor(mol.coords_bohr, device="cuda", requires_grad=True)
    J, K = direct_jk_autograd(ctx, D, R)
    E = 0.5 * (D * J).sum() - 0.25 * (D * K).sum()
    E.backward()
    Dd = D.detach()
    g2 = direct_jk_grad(ctx, Dd, 0.5 * Dd, -0.25 * Dd).cpu().numpy()
    got = R.grad.cpu().numpy()
    assert np.abs(got - g2).max() < 1e-10 * np.abs(g2).max()
    g = rhf_nuc_grad(res)
    assert np.abs(g.sum(axis=0)).max() < 1e-8
    print("[configs] (H2O)16 cc-pVDZ: central differences", flush=True)
    os.environ["ASUKA_DIRECT_JK_EPS_DENSITY"] = "0"
    try:
        kw = dict(two_e_backend="direct", eps_schwarz=1e-14, conv_tol=1e-12, conv_tol_dm=1e-9)
        rt = run_hf_df(mol, dm0=res.scf.dm, **kw)
        assert rt.scf.converged
        gt = rhf_nuc_grad(rt)
        picks = np.random.default_rng(4).choice(3 * mol.natm, 6, replace=False)[:2]
        R0 = np.asarray(mol.coords_bohr, dtype=np.float64)
        h = 1e-4
        for k in picks:
            a, x = divmod(int(k), 3)
            e = []
            for sgn in (-1, 1):
                Rk = R0.copy()
                Rk[a, x] += sgn * h
                r = run_hf_df(mol.with_coords(Rk), dm0=rt.scf.dm, **kw)
                assert r.scf.converged
                e.append(r.scf.e_tot)
            fd = (e[1] - e[0]) / (2 * h)
            print(f"[configs] atom {a} axis {x}: analytic {gt[a, x]:.9f} FD {fd:.9f}", flush=True)
            assert abs(fd - gt[a, x]) < 1e-6
    finally:
        del os.environ["ASUKA_DIRECT_JK_EPS_DENSITY"]
