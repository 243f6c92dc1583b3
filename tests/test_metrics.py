"""EFM-v1 counting convention, pinned by hand counts."""

from asuka_amd import metrics


def test_flop_eri_ssss_and_pppp():
    # (ss|ss), K = 1: n = 1, per root 24 + 15 + 6 + 4 = 49 -> 24 + 49 = 73
    assert metrics.flop_eri(0, 0, 0, 0, 1) == 73
    # (pp|pp), K = 2: L = 4, n = 3; (nmax+1)(mmax+1) = 9 -> 135; N1 = 16 -> 96; ncomp 81 -> 324
    per_root = 24 + 135 + 96 + 324
    assert metrics.flop_eri(1, 1, 1, 1, 2) == 2 * (24 + 3 * per_root)


def test_flop_dig_bins():
    # (ss|ss) with all flags set: jt = 2, kt = 4 -> 2 * 1 * 6 = 12 per quartet
    h = [0] * 8
    h[7] = 5
    assert metrics.flop_dig(0, 0, 0, 0, h) == 5 * 12
    # diagonal quartet of a same-shell pair: jt = 1, kt = 1 -> 4 per component
    h = [0] * 8
    h[0] = 1
    assert metrics.flop_dig(1, 0, 1, 0, h) == 2 * 9 * 2


def test_bytes_match_reference_cost_model_counts():
    # tests/direct/test_direct_jk_cost_model.py:30-63: 5 (ss|ss) tasks in bin 0
    # give density_loads == 10 (80 B); the warp model adds one J and one K
    # atomic per task (asuka/hf/direct_jk_cost.py:219-250) -> 10 atomics * 16 B.
    h = [0] * 8
    h[0] = 5
    assert metrics.bytes_alg(0, 0, 0, 0, h) == 80 + 160
    # bin 7 (all distinct): loads 6 -> 48 B; atomics J 2+2, K 2*(1+1+1+1)=8 -> 12*16 = 192 B
    h = [0] * 8
    h[7] = 1
    assert metrics.bytes_alg(0, 0, 0, 0, h) == 48 + 192


def test_build_work_sums():
    rows = [((0, 0, 0, 0), [1, 0, 0, 0, 0, 0, 0, 2], 10)]
    w = metrics.build_work(rows)
    assert w.quartets == 3
    assert w.prim_quartets == 10
    assert w.flop_eri == 10 * 73
    assert w.flop_dig == 1 * 4 + 2 * 12


def test_flop_grad_closed_forms():
    from asuka_amd import metrics

    # (ss|ss): one root, (nmax+2)(mmax+2) = 4, N1' = 16, one component
    assert metrics.flop_grad(0, 0, 0, 0, 1) == 24 + (24 + 15 * 4 + 6 * 16 + 48)
    # (pp|pp): n' = 3, 16 VRR entries, N1' = 81, 81 components, K = 3
    assert metrics.flop_grad(1, 1, 1, 1, 3) == 3 * (24 + 3 * (24 + 15 * 16 + 6 * 81 + 48 * 81))
    # linear in K
    assert metrics.flop_grad(2, 1, 1, 0, 7) == 7 * metrics.flop_grad(2, 1, 1, 0, 1)
