"""The Fock build captured in a HIP graph (torch.cuda.CUDAGraph) and replayed.

The build launches only kernels on the caller's stream (no memset, no host
synchronisation, no allocation), so it can be captured. Every replay is
checked against an eager build of the same density, with eager work and
synchronisations between replays, and with the density changed in place
between replays (the graph reads the captured buffer)."""

import numpy as np
import pytest

from asuka_amd.basis import build_ao_basis_cart
from asuka_amd.molecule import water_cluster

pytestmark = pytest.mark.gpu


def test_fock_build_graph_replay(gpu):
    from asuka_amd.hf.direct_jk import direct_fock_rhf, make_direct_jk_context

    torch = gpu
    mol = water_cluster(2, 3, "cc-pVTZ")
    ao, _ = build_ao_basis_cart(mol)
    ctx = make_direct_jk_context(ao, eps_schwarz=1e-12)
    nao = ao.nao
    rng = np.random.default_rng(5)

    def density(seed):
        C = np.random.default_rng(seed).standard_normal((nao, 10))
        C, _ = np.linalg.qr(C)
        return torch.tensor(2.0 * C @ C.T, device="cuda")

    h = torch.tensor(rng.standard_normal((nao, nao)), device="cuda")
    h = (h + h.T).contiguous()
    D = density(1)
    F = torch.empty((nao, nao), dtype=torch.float64, device="cuda")
    s = torch.cuda.Stream()
    s.wait_stream(torch.cuda.current_stream())
    with torch.cuda.stream(s):
        for _ in range(2):
            direct_fock_rhf(ctx, D, h, out=F)
    torch.cuda.current_stream().wait_stream(s)
    torch.cuda.synchronize()
    g = torch.cuda.CUDAGraph()
    with torch.cuda.graph(g):
        direct_fock_rhf(ctx, D, h, out=F)
    torch.cuda.synchronize()
    for rep, seed in enumerate([1, 2, 2, 3]):
        D.copy_(density(seed))
        torch.cuda.synchronize()
        ref = direct_fock_rhf(ctx, D.clone(), h)
        torch.cuda.synchronize()
        g.replay()
        torch.cuda.synchronize()
        err = float((F - ref).abs().max())
        assert err < 1e-11, (rep, err)


def test_batched_build_matches_single_and_replays(gpu):
    """Batched J/K over molecules of different sizes equals the per-molecule
    builds exactly (same kernels, same order), eagerly and as a replayed graph."""
    from asuka_amd.hf.direct_jk import direct_JK, direct_JK_batched, make_direct_jk_context

    torch = gpu
    ctxs, Ds = [], []
    for n, seed in [(1, 10), (3, 11), (2, 12)]:
        mol = water_cluster(n, seed, "cc-pVDZ")
        ao, _ = build_ao_basis_cart(mol)
        ctxs.append(make_direct_jk_context(ao, eps_schwarz=1e-12))
        C = np.random.default_rng(seed).standard_normal((ao.nao, 5 * n))
        C, _ = np.linalg.qr(C)
        Ds.append(torch.tensor(2.0 * C @ C.T, device="cuda"))
    ref = [direct_JK(c, D) for c, D in zip(ctxs, Ds)]
    out = direct_JK_batched(ctxs, Ds)
    for (J0, K0), (J1, K1) in zip(ref, out):
        assert float((J0 - J1).abs().max()) < 1e-12
        assert float((K0 - K1).abs().max()) < 1e-12
    s = torch.cuda.Stream()
    s.wait_stream(torch.cuda.current_stream())
    with torch.cuda.stream(s):
        direct_JK_batched(ctxs, Ds, out=out)
    torch.cuda.current_stream().wait_stream(s)
    torch.cuda.synchronize()
    g = torch.cuda.CUDAGraph()
    with torch.cuda.graph(g):
        direct_JK_batched(ctxs, Ds, out=out)
    for rep in range(3):
        for J, K in out:
            J.zero_()
            K.zero_()
        torch.cuda.synchronize()
        g.replay()
        torch.cuda.synchronize()
        for (J0, K0), (J1, K1) in zip(ref, out):
            assert float((J0 - J1).abs().max()) < 1e-12, rep
            assert float((K0 - K1).abs().max()) < 1e-12, rep
