"""Batch contexts (asuka_jk_batch_create / asuka_jk_batch_build): several
independent molecules in one context, one launch per class pair for the whole
batch, each molecule's D/J/K a packed block.

* against the oracle's golden J/K of (H2O)2 cc-pVTZ and (H2O)4 cc-pVDZ,
  batched together with a water monomer (1e-10 absolute);
* against one-molecule contexts (1e-12), eagerly and replayed from a HIP
  graph three times;
* a tuning table from one molecule applies to the batch; single-molecule entry
  points reject a batch context.
"""

import ctypes
import os

import numpy as np
import pytest

from asuka_amd.basis import build_ao_basis_cart
from asuka_amd.molecule import water_cluster, water_monomer

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))


@pytest.fixture(scope="module")
def jk_golden():
    return dict(np.load(os.path.join(HERE, "golden", "jk_golden.npz"), allow_pickle=False))


def _density(nao, nocc, seed):
    C = np.random.default_rng(seed).standard_normal((nao, nocc))
    C, _ = np.linalg.qr(C)
    return C @ C.T


def test_batch_against_golden_and_single(gpu, jk_golden):
    from asuka_amd.hf.direct_jk import direct_JK, direct_JK_batched, make_direct_jk_batch_context, \
        make_direct_jk_context

    mols = [water_cluster(2, 2, "cc-pVTZ"), water_monomer("cc-pVDZ"), water_cluster(4, 4, "cc-pVDZ")]
    aos = [build_ao_basis_cart(m)[0] for m in mols]
    bc = make_direct_jk_batch_context(aos, eps_schwarz=1e-12)
    assert bc.naos == [a.nao for a in aos]
    Ds = [jk_golden["h2o2_tz_D"], _density(aos[1].nao, 5, 3), jk_golden["h2o4_dz_D"]]
    out = direct_JK_batched(bc, Ds)
    for (J, K), case in ((out[0], "h2o2_tz"), (out[2], "h2o4_dz")):
        assert np.max(np.abs(J.cpu().numpy() - jk_golden[f"{case}_J_scr"])) < 1e-10
        assert np.max(np.abs(K.cpu().numpy() - jk_golden[f"{case}_K_scr"])) < 1e-10
    refs = [direct_JK(make_direct_jk_context(a, eps_schwarz=1e-12), D) for a, D in zip(aos, Ds)]
    for (J, K), (J0, K0) in zip(out, refs):
        assert float((J - J0).abs().max()) < 1e-12 and float((K - K0).abs().max()) < 1e-12
        assert float((J - J.T).abs().max()) == 0.0
    # J only, into caller buffers that hold garbage
    torch = gpu
    bufs = [(torch.full_like(J, 3.0), None) for J, _ in out]
    got = direct_JK_batched(bc, Ds, want_K=False, out=bufs)
    for (J, K), (J0, _) in zip(got, out):
        assert K is None and float((J - J0).abs().max()) < 1e-13


def test_batch_graph_replay_and_tuning(gpu):
    from asuka_amd.hf.direct_jk import direct_JK_batched, make_direct_jk_batch_context, make_direct_jk_context

    torch = gpu
    mols = [water_cluster(n, 10 + n, "cc-pVDZ") for n in (3, 1, 2)]
    aos = [build_ao_basis_cart(m)[0] for m in mols]
    Ds = [torch.tensor(_density(a.nao, 5 * n, 20 + n), device="cuda") for a, n in zip(aos, (3, 1, 2))]
    bc = make_direct_jk_batch_context(aos)
    ref = [(J.clone(), K.clone()) for J, K in direct_JK_batched(bc, Ds)]
    out = direct_JK_batched(bc, Ds)
    s = torch.cuda.Stream()
    s.wait_stream(torch.cuda.current_stream())
    with torch.cuda.stream(s):
        direct_JK_batched(bc, Ds, out=out)
    torch.cuda.current_stream().wait_stream(s)
    torch.cuda.synchronize()
    g = torch.cuda.CUDAGraph()
    with torch.cuda.graph(g):
        direct_JK_batched(bc, Ds, out=out)
    for rep in range(3):
        for J, K in out:
            J.zero_()
            K.zero_()
        torch.cuda.synchronize()
        g.replay()
        torch.cuda.synchronize()
        for (J, K), (J0, K0) in zip(out, ref):
            assert float((J - J0).abs().max()) < 1e-12, rep
            assert float((K - K0).abs().max()) < 1e-12, rep
    # a table tuned on one molecule applies to the batch and keeps the results
    one = make_direct_jk_context(aos[0])
    table = one.autotune(Ds[0], reps=1)
    bc.set_tuning(table)
    for (J, K), (J0, K0) in zip(direct_JK_batched(bc, Ds), ref):
        assert float((J - J0).abs().max()) < 1e-12 and float((K - K0).abs().max()) < 1e-12
    # single-molecule entry points reject a batch context
    from asuka_amd import _lib

    J = torch.empty((aos[0].nao, aos[0].nao), dtype=torch.float64, device="cuda")
    with pytest.raises(ValueError, match="batch"):
        _lib.check(_lib.lib().asuka_jk_build(ctypes.c_void_p(bc.handle), _lib.ptr(Ds[0]), _lib.ptr(J), None,
                                             _lib.WANT_J, ctypes.c_void_p(_lib.stream_handle())))
