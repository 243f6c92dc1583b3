#!/usr/bin/env python
"""Benchmark of the integral-direct J/K engine on one MI355X.

Default workload: (H2O)64 water cluster (4x4x4 grid, 3.0 A, seed 5), cc-pVDZ
cartesian (1600 AOs), D = 2 C C^T with C = QR(N(0,1) 1600 x 320) (seed 1005),
one direct RHF Fock build F = h + J - K/2 per step at Schwarz 1e-12 and density
screening 1e-10, timed with the all-FP64 build and with the mixed FP32/FP64
build. --config 2 times one J/K build of (H2O)8 cc-pVTZ (520 AOs); --config 3
the batched build of 32 clusters; --config 4 the (H2O)16 cc-pVDZ gradient;
--config 5 --scf 15 RHF iterations.

Prints one JSON line: shell quartets/s of the mixed build (value), seconds per
build in both modes and their accuracy (max |dF|, max |dJ|, max |dK|, dE), the
EFM-v1 roofline of the ERI/digestion kernels (HIP events on the launch
stream), a parity stamp against the CPU oracle on sampled tiles, and the CPU
baseline (the oracle's reference-arithmetic build on a bounded sample of the
same screened quartets, timed on this host's cores).

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--config 2|3|4|5] [--mode fp64|mixed|both]
"""

from __future__ import annotations

import argparse
import ctypes
import json
import os
import platform
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "Fock-build s & quartets/s: mixed vs FP64 GPU at ≤1e-6 Eh; gradient s/eval"

# HBM bytes of one build's ERI kernels (FETCH_SIZE + WRITE_SIZE, each scaled
# by what the counter reads for a 1 GiB device copy), measured with rocprofv3
# --pmc per workload and mode (profiles/*_hbm_traffic.md); None where not
# measured.
HBM_TRAFFIC = {(2, "fp64"): 1.247e10, (2, "mixed"): 1.625e10, (5, "fp64"): 5.505e11, (5, "mixed"): 5.176e11}

CONFIGS = {
    2: dict(n=8, seed=2, basis="cc-pVTZ", nocc=40, cseed=1002, eps=1e-12, kind="jk",
            workload="(H2O)8 cc-pVTZ cart, one direct J/K build"),
    5: dict(n=64, seed=5, basis="cc-pVDZ", nocc=320, cseed=1005, eps=1e-12, kind="fock",
            workload="(H2O)64 cc-pVDZ cart, one direct RHF Fock build",
            workload_scf="(H2O)64 cc-pVDZ cart, 15 direct RHF iterations"),
    3: dict(nmol=32, seed=3, basis="cc-pVDZ", eps=1e-12,
            workload="32 water clusters of 1-16 molecules, cc-pVDZ cart, one batched J/K build"),
    4: dict(n=16, seed=4, basis="cc-pVDZ", eps=1e-12,
            workload="(H2O)16 cc-pVDZ cart, converged RHF D: forward J/K(D,R) + backward dE/dR"),
}


def run_batched(args, cfg) -> int:
    """Batched workload: 32 independent clusters (sizes default_rng(3).integers(1,
    17, 32), placement seeds 1000 + i), random orthonormal C with 5 orbitals per
    water (seed 2000 + i), one batched J/K build per step on a batch context
    (one launch per class pair for the whole batch), eager and replayed from a
    captured HIP graph, with the default launch shapes and with the shapes the
    autotuner picks on the largest molecule. The per-molecule-context path
    (builds issued in turn) is timed beside it."""
    import torch

    from asuka_amd.basis import build_ao_basis_cart
    from asuka_amd.hf.direct_jk import direct_JK_batched, make_direct_jk_batch_context, make_direct_jk_context
    from asuka_amd.molecule import water_cluster

    dev = torch.device("cuda", 0)
    torch.cuda.set_device(dev)
    sizes = np.random.default_rng(cfg["seed"]).integers(1, 17, cfg["nmol"])
    aos, Ds = [], []
    for i, n in enumerate(sizes):
        ao, _ = build_ao_basis_cart(water_cluster(int(n), 1000 + i, cfg["basis"]))
        aos.append(ao)
        C = np.random.default_rng(2000 + i).standard_normal((ao.nao, 5 * int(n)))
        C, _ = np.linalg.qr(C)
        Ds.append(torch.tensor(2.0 * C @ C.T, device=dev))
    t0 = time.perf_counter()
    bc = make_direct_jk_batch_context(aos, eps_schwarz=cfg["eps"])
    t_ctx = time.perf_counter() - t0
    t0 = time.perf_counter()
    ctxs = [make_direct_jk_context(ao, eps_schwarz=cfg["eps"]) for ao in aos]
    t_ctxs = time.perf_counter() - t0
    nq = bc.ntasks
    log(f"batch: {len(aos)} molecules, {sum(bc.naos)} AOs, {nq} Schwarz-screened quartets "
        f"(batch context {t_ctx:.1f}s, per-molecule contexts {t_ctxs:.1f}s)")

    def measure(target):
        out = direct_JK_batched(target, Ds)
        for _ in range(max(0, args.warmup)):
            direct_JK_batched(target, Ds, out=out)
        torch.cuda.synchronize(dev)
        t0 = time.perf_counter()
        for _ in range(args.steps):
            direct_JK_batched(target, Ds, out=out)
        torch.cuda.synchronize(dev)
        s_eager = (time.perf_counter() - t0) / args.steps
        side = torch.cuda.Stream()
        side.wait_stream(torch.cuda.current_stream())
        with torch.cuda.stream(side):
            direct_JK_batched(target, Ds, out=out)
        torch.cuda.current_stream().wait_stream(side)
        torch.cuda.synchronize(dev)
        graph = torch.cuda.CUDAGraph()
        with torch.cuda.graph(graph):
            direct_JK_batched(target, Ds, out=out)
        graph.replay()
        torch.cuda.synchronize(dev)
        t0 = time.perf_counter()
        for _ in range(args.steps):
            graph.replay()
        torch.cuda.synchronize(dev)
        return s_eager, (time.perf_counter() - t0) / args.steps, [x.clone() for p in out for x in p if x is not None]

    s_eager_m, s_graph_m, ref = measure(ctxs)
    log(f"per-molecule contexts: eager {s_eager_m * 1e3:.2f} ms, graph {s_graph_m * 1e3:.2f} ms")
    s_eager0, s_graph0, got0 = measure(bc)
    d0 = max(float((a - b).abs().max()) for a, b in zip(got0, ref))
    log(f"batch context, default shapes: eager {s_eager0 * 1e3:.2f} ms, graph {s_graph0 * 1e3:.2f} ms "
        f"(max |dJ|,|dK| vs per-molecule contexts {d0:.1e})")
    # autotune on the largest molecule, apply its table to the batch
    big = int(np.argmax([a.nao for a in aos]))
    t0 = time.perf_counter()
    table = ctxs[big].autotune(Ds[big])
    bc.set_tuning(table)
    t_tune = time.perf_counter() - t0
    s_eager, s_graph, got = measure(bc)
    dmax = max(float((a - b).abs().max()) for a, b in zip(got, ref))
    log(f"batch context, autotuned shapes: eager {s_eager * 1e3:.2f} ms, graph {s_graph * 1e3:.2f} ms "
        f"(tuning {t_tune:.2f}s, max |dJ|,|dK| vs per-molecule contexts {dmax:.1e})")
    res = {
        "metric": "batched J/K build: Schwarz-screened shell quartets/s over the batch (graph replay)",
        "value": nq / s_graph,
        "unit": "shell_quartets/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": s_graph * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (seeded water clusters, random orthonormal occupied orbitals)",
        "config": {"workload": cfg["workload"], "sizes": [int(n) for n in sizes], "basis": cfg["basis"],
                   "eps_schwarz": cfg["eps"], "launches": "one per class pair for the whole batch",
                   "launch_shapes": "autotuned"},
        "s_per_batch_eager": s_eager,
        "s_per_batch_graph": s_graph,
        "graph_speedup": s_eager / s_graph,
        "default_shapes": {"s_per_batch_eager": s_eager0, "s_per_batch_graph": s_graph0},
        "autotuned_vs_default": s_graph0 / s_graph,
        "autotune_s": t_tune,
        "autotune_table_from": f"molecule {big} ({aos[big].nao} AOs), applied to the batch",
        "per_molecule_contexts": {"s_per_batch_eager": s_eager_m, "s_per_batch_graph": s_graph_m},
        "batch_context_vs_per_molecule": s_graph_m / s_graph,
        "max_abs_dJK_vs_per_molecule_contexts": max(d0, dmax),
        "context_build_s": t_ctx,
    }
    print(json.dumps(res), flush=True)
    return 0


def run_scf(args, cfg) -> int:
    """Large direct SCF: (H2O)64 cc-pVDZ (1600 AOs), 15 RHF iterations (core
    guess, one predictor Fock build, DIIS) with the all-FP64 build and with the
    mixed FP32/FP64 build, on one context, each SCF iteration one replay of a
    captured HIP graph (Fock build, device DIIS, purified density, energy and
    convergence test); the energy after 15 iterations of each is compared. The
    time outside the Fock builds is the SCF time minus that of the same number
    of Fock builds alone on the same densities: the SCF keeps the input density
    of every Fock build it makes (the predictor's and each cycle's), and each is
    built again alone, timed with a device synchronisation on both sides (the
    work of a build depends on its density through the density screening and
    the FP32 routing, so builds at another density would misstate it)."""
    import torch

    from asuka_amd.frontend import build_int1e
    from asuka_amd.basis import build_ao_basis_cart
    from asuka_amd.hf.direct_jk import make_direct_jk_context
    from asuka_amd.hf.direct_scf import rhf_direct
    from asuka_amd.molecule import water_cluster

    dev = torch.device("cuda", 0)
    torch.cuda.set_device(dev)
    mol = water_cluster(cfg["n"], cfg["seed"], cfg["basis"])
    ao, _ = build_ao_basis_cart(mol)
    t0 = time.perf_counter()
    ctx = make_direct_jk_context(ao, eps_schwarz=cfg["eps"], tau_mixed=args.tau, atom_coords_bohr=mol.coords_bohr,
                                 atom_charges=mol.charges)
    t_ctx = time.perf_counter() - t0
    ints = build_int1e(ctx)
    niter = 15
    out = {}
    kw = dict(nelec=mol.nelectron, enuc=float(mol.energy_nuc()))
    D15 = None
    from asuka_amd.hf.direct_jk import direct_fock_rhf

    hd = ints.hcore if isinstance(ints.hcore, torch.Tensor) else torch.as_tensor(np.asarray(ints.hcore), device=dev)
    for mode in ("fp64", "mixed"):
        mixed = mode == "mixed"
        # warm-up: one short SCF (graph capture, allocator pools)
        rhf_direct(ints.S, ints.hcore, ctx, max_cycle=2, conv_tol=0.0, conv_tol_dm=0.0, mixed_precision=mixed, **kw)
        torch.cuda.synchronize(dev)
        prof = {"keep_densities": True}
        t0 = time.perf_counter()
        r = rhf_direct(ints.S, ints.hcore, ctx, max_cycle=niter, conv_tol=0.0, conv_tol_dm=0.0,
                       mixed_precision=mixed, profile=prof, **kw)
        torch.cuda.synchronize(dev)
        t_scf = time.perf_counter() - t0
        d_cyc = prof["scf"].pop("cycle_densities", [])
        d_pre = prof["scf"].pop("predictor_densities", [])
        F = torch.empty_like(hd)

        def fock_alone(dens):
            t = 0.0
            for Dk in dens:
                torch.cuda.synchronize(dev)
                t1 = time.perf_counter()
                direct_fock_rhf(ctx, Dk, hd, mixed_precision=mixed, out=F)
                torch.cuda.synchronize(dev)
                t += time.perf_counter() - t1
            return t

        t_fock_loop = fock_alone(d_cyc)  # the replayed cycles' builds alone
        t_fock = t_fock_loop + fock_alone(d_pre)  # every build of the SCF alone
        nfock = len(d_cyc) + len(d_pre)
        del d_cyc, d_pre
        nrep = int(prof["scf"].get("cycles_replayed", 0))
        t_loop = float(prof["scf"].get("cycle_loop_s", 0.0))
        out[mode] = {"s": t_scf, "e_tot": r.e_tot, "iters": r.niter, "fock_builds": nfock, "fock_s": t_fock,
                     "non_fock_fraction": (t_scf - t_fock) / t_scf, "graph": bool(prof["scf"].get("graph")),
                     "cycles_replayed": nrep, "cycle_loop_s": t_loop,
                     "non_fock_fraction_cycles": (t_loop - t_fock_loop) / t_loop if t_loop > 0 else None,
                     "purify_steps": prof["scf"].get("purify_steps"),
                     "purify_err_max": prof["scf"].get("purify_err_max")}
        if mode == "fp64":
            D15 = r.dm
        log(f"{mode}: {niter} SCF iterations in {t_scf:.2f} s ({nfock} Fock builds alone {t_fock:.2f} s, "
            f"same densities, outside the builds {100 * (t_scf - t_fock) / t_scf:.1f} %; the {nrep} graph-replayed cycles "
            f"{t_loop:.2f} s, outside their builds {100 * (t_loop - t_fock_loop) / max(t_loop, 1e-9):.1f} %), "
            f"E = {r.e_tot:.10f}")
    dE = out["mixed"]["e_tot"] - out["fp64"]["e_tot"]
    # precision of one build at the FP64 run's iteration-15 density:
    # E = 1/2 sum D (h + F), so dE = 1/2 sum D (F_mixed - F_fp64)
    F64 = direct_fock_rhf(ctx, D15, hd, mixed_precision=False)
    F32 = direct_fock_rhf(ctx, D15, hd, mixed_precision=True)
    dE_build = float(0.5 * (D15 * (F32 - F64)).sum())
    # converged energies, both modes continued from that density (the mixed
    # SCF builds in FP64 once ||D_new - D||_F < 1e-3, the driver's default)
    conv = {}
    for mode in ("fp64", "mixed"):
        pc = {}
        r = rhf_direct(ints.S, ints.hcore, ctx, max_cycle=60, conv_tol=1e-9, conv_tol_dm=1e-6, dm0=D15,
                       mixed_precision=(mode == "mixed"), profile=pc, **kw)
        conv[mode] = {"e_tot": r.e_tot, "iters": r.niter, "converged": bool(r.converged),
                      "fp64_builds_from_cycle": pc["scf"].get("fp64_from_cycle")}
        log(f"{mode}: converged={r.converged} after {r.niter} more iterations, E = {r.e_tot:.10f}"
            + (f" (FP64 builds from cycle {conv[mode]['fp64_builds_from_cycle']})" if mode == "mixed" else ""))
    dE_conv = conv["mixed"]["e_tot"] - conv["fp64"]["e_tot"]
    res = {
        "metric": "direct RHF: seconds for 15 SCF iterations, mixed vs FP64 GPU build",
        "value": out["mixed"]["s"],
        "unit": "s",
        "n_gpus": 1,
        "steps": 1,
        "warmup": 0,
        "ms_per_step": out["mixed"]["s"] * 1e3,
        "higher_is_better": False,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64+f32",
        "data": "synthetic (seeded water cluster geometry)",
        "config": {"workload": cfg["workload_scf"], "nao": ao.nao, "basis": cfg["basis"], "n_water": cfg["n"],
                   "geometry_seed": cfg["seed"], "eps_schwarz": cfg["eps"], "iterations": niter,
                   "tau_mixed": args.tau},
        "fp64_s": out["fp64"]["s"],
        "mixed_s": out["mixed"]["s"],
        "speedup_mixed_vs_fp64": out["fp64"]["s"] / out["mixed"]["s"],
        "per_mode": out,
        "e_tot_fp64": out["fp64"]["e_tot"],
        "e_tot_mixed": out["mixed"]["e_tot"],
        "dE_mixed_vs_fp64_after_15_iterations_Eh": dE,
        "dE_mixed_vs_fp64_one_build_at_fp64_density_Eh": dE_build,
        "dE_mixed_vs_fp64_converged_Eh": dE_conv,
        "converged": conv,
        "fock_builds_per_scf": niter + 1,
        "context_build_s": t_ctx,
    }
    print(json.dumps(res), flush=True)
    return 0


def run_gradient(args, cfg) -> int:
    """Gradient workload: one step = forward J/K through the autograd Function
    at the converged RHF density, E_2e = 1/2 Tr(D J) - 1/4 Tr(D K), and the
    hand-written backward to dE_2e/dR (plus dE/dD), FP64."""
    import torch

    from asuka_amd.frontend import run_hf_df
    from asuka_amd.hf.grad import direct_jk_autograd, rhf_nuc_grad
    from asuka_amd.molecule import water_cluster

    dev = torch.device("cuda", 0)
    torch.cuda.set_device(dev)
    mol = water_cluster(cfg["n"], cfg["seed"], cfg["basis"])
    t0 = time.perf_counter()
    res = run_hf_df(mol, two_e_backend="direct", eps_schwarz=cfg["eps"], conv_tol=1e-10, mixed_precision=False)
    t_scf = time.perf_counter() - t0
    ctx = res.direct_jk_ctx
    log(f"SCF: nao={ctx.nao} E={res.scf.e_tot:.10f} converged={res.scf.converged} in {t_scf:.1f}s")
    D0 = res.scf.dm.detach().clone()
    R0 = torch.tensor(mol.coords_bohr, device=dev)

    def step(mixed):
        D = D0.clone().requires_grad_(True)
        R = R0.clone().requires_grad_(True)
        J, K = direct_jk_autograd(ctx, D, R, mixed_precision=mixed)
        E = 0.5 * (D * J).sum() - 0.25 * (D * K).sum()
        E.backward()
        return R.grad, D.grad, float(E)

    def timed(mixed):
        for _ in range(max(0, args.warmup)):
            step(mixed)
        torch.cuda.synchronize(dev)
        t0 = time.perf_counter()
        for _ in range(args.steps):
            r = step(mixed)
        torch.cuda.synchronize(dev)
        return (time.perf_counter() - t0) / args.steps, r

    s_per, (g2, gD, E64) = timed(False)
    s_mix, (g2m, gDm, E32) = timed(True)
    mixed = {"s_per_eval": s_mix, "speedup_vs_fp64": s_per / s_mix, "dE_2e_Eh": E32 - E64,
             "max_abs_dgrad_R": float((g2m - g2).abs().max()), "max_abs_dgrad_D": float((gDm - gD).abs().max())}
    log(f"gradient step: FP64 {s_per * 1e3:.1f} ms, mixed forward and dL/dD {s_mix * 1e3:.1f} ms "
        f"(dE {E32 - E64:.1e}, max |dgrad_R| {mixed['max_abs_dgrad_R']:.1e})")
    t0 = time.perf_counter()
    g = rhf_nuc_grad(res)
    torch.cuda.synchronize(dev)
    t_full = time.perf_counter() - t0

    # roofline of the derivative kernels (asuka_jk_grad, the step's dominant
    # launch set): EFM-v1 FLOP_grad over the Schwarz-screened quartets (the
    # gradient applies no density screening) / their event-timed duration on
    # the stream they run on
    from asuka_amd import metrics
    from asuka_amd.hf.grad import direct_jk_grad

    stats = ctx.class_stats()
    flop_grad = sum(metrics.flop_grad(*r["class"], r["prim_quartets0"]) for r in stats)
    stream = torch.cuda.current_stream(dev)
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    direct_jk_grad(ctx, D0, 0.5 * D0, -0.25 * D0)
    torch.cuda.synchronize(dev)
    reps = max(1, args.steps)
    e0.record(stream)
    for _ in range(reps):
        direct_jk_grad(ctx, D0, 0.5 * D0, -0.25 * D0)
    e1.record(stream)
    torch.cuda.synchronize(dev)
    t_gk = e0.elapsed_time(e1) / 1e3 / reps
    achieved = flop_grad / t_gk / 1e12
    roof = {"bound": "valu", "achieved": achieved, "peak": metrics.PEAK_FP64_TFLOPS, "unit": "TFLOP/s",
            "frac": achieved / metrics.PEAK_FP64_TFLOPS, "traffic": None,
            "flops_grad_efm_v1": flop_grad, "grad_kernel_ms": t_gk * 1e3,
            "schwarz_quartets": int(sum(r["quartets"] for r in stats))}
    log(f"derivative kernels: {t_gk * 1e3:.2f} ms, EFM-v1 FLOP_grad {flop_grad / 1e9:.1f} G, {achieved:.2f} TFLOP/s")
    cpu = None
    if not args.no_cpu:
        try:  # the oracle's derivative contraction on a bounded sample of the quartets
            from oracle import oracle as O

            from asuka_amd.basis import build_ao_basis_cart

            threads = len(os.sched_getaffinity(0))
            ao, _ = build_ao_basis_cart(mol)
            octx = O.make_context(ao)
            Dh = D0.cpu().numpy()
            Q = ctx.Q_sp
            ones = np.ones_like(Q)
            rng = np.random.default_rng(11)
            ab, cd = sample_screened(Q, ones, cfg["eps"], 0.0, 2000, rng)
            sa = np.asarray(ao.shell_atom, dtype=np.int32)
            t1 = time.perf_counter()
            O.grad2e_tasks(octx, Dh, 0.5 * Dh, -0.25 * Dh, sa, mol.natm, ab, cd, hp=False, threads=threads)
            rate0 = ab.size / (time.perf_counter() - t1)
            nq = roof["schwarz_quartets"]
            nsamp = int(min(nq, max(2000, rate0 * args.cpu_seconds)))
            ab, cd = sample_screened(Q, ones, cfg["eps"], 0.0, nsamp, rng)
            t1 = time.perf_counter()
            O.grad2e_tasks(octx, Dh, 0.5 * Dh, -0.25 * Dh, sa, mol.natm, ab, cd, hp=False, threads=threads)
            tc = time.perf_counter() - t1
            rate = ab.size / tc
            cpu = {"value": rate / nq, "unit": "gradient_evals/s (derivative contraction, extrapolated)",
                   "cores": threads, "kind": "port",
                   "sample": f"{ab.size} of the {nq} Schwarz-screened quartets drawn uniformly, the oracle's "
                             f"adjoint tile and derivative contraction per quartet, {tc:.1f} s on {cpu_model()}",
                   "quartets_per_s": rate}
        except Exception as e:  # the baseline is reported, never required
            log("oracle leg skipped:", repr(e))
    fd = None
    if args.fd:
        # central differences (step 1e-4 Bohr) of the converged SCF energy on 6
        # coordinates drawn with default_rng(seed).choice(3 natm, 6). Screening
        # is tightened (Schwarz 1e-14, no density screening) for the check, so
        # the screened-away set does not jump between displaced geometries.
        os.environ["ASUKA_DIRECT_JK_EPS_DENSITY"] = "0"
        kw = dict(two_e_backend="direct", eps_schwarz=1e-14, conv_tol=1e-12, conv_tol_dm=1e-9, mixed_precision=False)
        rt = run_hf_df(mol, dm0=res.scf.dm, **kw)
        gt = rhf_nuc_grad(rt)
        h = 1e-4
        picks = np.random.default_rng(cfg["seed"]).choice(3 * mol.natm, 6, replace=False)
        R = np.asarray(mol.coords_bohr, dtype=np.float64)
        worst = 0.0
        for k in picks:
            a, x = divmod(int(k), 3)
            e = []
            for sgn in (-1, 1):
                Rk = R.copy()
                Rk[a, x] += sgn * h
                r = run_hf_df(mol.with_coords(Rk), dm0=rt.scf.dm, **kw)
                e.append(r.scf.e_tot)
            fdv = (e[1] - e[0]) / (2 * h)
            worst = max(worst, abs(fdv - float(gt[a, x])))
            log(f"FD atom {a} axis {x}: analytic {gt[a, x]:.9f} FD {fdv:.9f}")
        del os.environ["ASUKA_DIRECT_JK_EPS_DENSITY"]
        fd = {"coordinates": [int(k) for k in picks], "step_bohr": h, "eps_schwarz": 1e-14, "eps_density": 0.0,
              "max_abs_diff_eh_per_bohr": worst,
              "max_abs_diff_vs_default_screening": float(np.abs(gt - g).max())}
    out = {
        "metric": "gradient s/eval (forward J/K + backward dE/dR and dE/dD, FP64)",
        "value": 1.0 / s_per,
        "unit": "gradient_evals/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": s_per * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (seeded water cluster geometry), converged RHF density",
        "config": {"workload": cfg["workload"], "nao": ctx.nao, "basis": cfg["basis"], "n_water": cfg["n"],
                   "geometry_seed": cfg["seed"], "eps_schwarz": cfg["eps"]},
        "s_per_eval": s_per,
        "full_rhf_gradient_s": t_full,
        "mixed": mixed,
        "max_abs_translational_sum": float(abs(g.sum(axis=0)).max()),
        "e_tot": res.scf.e_tot,
        "grad_2e_norm": float(g2.norm()),
        "fd_check": fd,
        "roofline": roof,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)
    return 0


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def cpu_model() -> str:
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return platform.processor() or "unknown"


def sample_screened(Q, Dsp, eps_s, eps_d, count, rng):
    """Uniform sample of the canonical screened quartet set: draw canonical
    (ab >= cd) pairs uniformly, keep those passing Schwarz and density screening."""
    nsp = Q.size
    out_ab, out_cd = [], []
    need = count
    while need > 0:
        m = max(4 * need, 100000)
        i = rng.integers(0, nsp, m)
        j = rng.integers(0, nsp, m)
        ab = np.maximum(i, j)
        cd = np.minimum(i, j)
        keep = (Q[ab] * Q[cd] >= eps_s) & ((Q[ab] * Q[cd]) * np.maximum(Dsp[ab], Dsp[cd]) >= eps_d) & (Q[ab] > 0) & (Q[cd] > 0)
        # off-diagonal pairs are drawn twice as often as diagonal ones: thin the diagonal-free mass
        diag = ab == cd
        keep &= diag | (rng.random(m) < 0.5)
        ab, cd = ab[keep][:need], cd[keep][:need]
        out_ab.append(ab)
        out_cd.append(cd)
        need -= ab.size
    return np.concatenate(out_ab).astype(np.int32), np.concatenate(out_cd).astype(np.int32)


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--config", type=int, default=5, choices=sorted(CONFIGS))
    ap.add_argument("--mode", default="both", choices=["fp64", "mixed", "both"])
    ap.add_argument("--cpu-seconds", type=float, default=15.0, help="target CPU-baseline sample time")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--fd", action="store_true", help="gradient workload: central-difference check on 6 coordinates")
    ap.add_argument("--autotune", action="store_true", help="J/K workloads: autotune launch shapes before timing")
    ap.add_argument("--scf", action="store_true", help="config 5: time 15 SCF iterations, FP64 and mixed")
    ap.add_argument("--tau", type=float, default=1e-5, help="FP32 cut on Q_ab Q_cd max|D| of the mixed build")
    ap.add_argument("--copy-probe", type=int, default=0, metavar="BYTES",
                    help="only run one device-to-device copy of BYTES (the calibration kernel of the HBM counters)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this engine runs on one GPU; multi-GPU sharding is design-only "
                                   "(one RCCL all-reduce of J/K per build), --gpus must be 1"}))
        return 2

    import torch

    if args.copy_probe > 0:
        # a plain copy of known size: reads and writes copy_probe bytes once
        n = args.copy_probe // 8
        src = torch.ones(n, dtype=torch.float64, device="cuda")
        dst = torch.empty_like(src)
        torch.cuda.synchronize()
        dst.copy_(src)
        torch.cuda.synchronize()
        print(json.dumps({"copy_probe_bytes": n * 8}), flush=True)
        return 0

    from asuka_amd import _lib, metrics
    from asuka_amd.basis import build_ao_basis_cart
    from asuka_amd.hf.direct_jk import eri_tiles, make_direct_jk_context
    from asuka_amd.molecule import water_cluster

    cfg = CONFIGS[args.config]
    if args.config == 4:
        return run_gradient(args, cfg)
    if args.config == 3:
        return run_batched(args, cfg)
    if args.config == 5 and args.scf:
        return run_scf(args, cfg)
    dev = torch.device("cuda", 0)
    torch.cuda.set_device(dev)
    mol = water_cluster(cfg["n"], cfg["seed"], cfg["basis"])
    ao, _ = build_ao_basis_cart(mol)
    nao = ao.nao
    C = np.random.default_rng(cfg["cseed"]).standard_normal((nao, cfg["nocc"]))
    C, _ = np.linalg.qr(C)
    fock = cfg["kind"] == "fock"
    D_h = (2.0 if fock else 1.0) * (C @ C.T)
    t0 = time.perf_counter()
    ctx = make_direct_jk_context(ao, eps_schwarz=cfg["eps"], tau_mixed=args.tau, atom_coords_bohr=mol.coords_bohr,
                                 atom_charges=mol.charges)
    t_ctx = time.perf_counter() - t0
    log(f"context: nao={nao} nsp={ctx.nsp} schwarz-screened quartets={ctx.ntasks} ({t_ctx:.2f}s)")

    D = torch.as_tensor(D_h, device=dev)
    hcore = None
    if fock:
        from asuka_amd.frontend import build_int1e

        hcore = build_int1e(ctx).hcore
        if not isinstance(hcore, torch.Tensor):
            hcore = torch.as_tensor(np.asarray(hcore), device=dev)
        hcore = hcore.to(device=dev, dtype=torch.float64).contiguous()
    t_tune = None
    if args.autotune:
        t0 = time.perf_counter()
        ctx.autotune(D)
        t_tune = time.perf_counter() - t0
        log(f"autotune: {t_tune:.2f}s")
    L = _lib.lib()
    h = ctypes.c_void_p(ctx.handle)
    stream = torch.cuda.current_stream(dev)
    sptr = ctypes.c_void_p(stream.cuda_stream)

    def run_mode(mode: str):
        J = torch.empty((nao, nao), dtype=torch.float64, device=dev)
        K = torch.empty_like(J)
        F = torch.empty_like(J)
        flags = _lib.WANT_J | _lib.WANT_K | (_lib.MIXED if mode == "mixed" else 0)

        def build(extra=0):
            if fock:
                _lib.check(L.asuka_fock_build(h, _lib.ptr(D), _lib.ptr(hcore), _lib.ptr(F),
                                              (flags & _lib.MIXED) | extra, sptr))
            else:
                _lib.check(L.asuka_jk_build(h, _lib.ptr(D), _lib.ptr(J), _lib.ptr(K), flags | extra, sptr))

        for _ in range(max(0, args.warmup)):
            build()
        torch.cuda.synchronize(dev)
        build(_lib.COUNT)  # counted build: exact per-class work of this density
        stats = ctx.class_stats()
        ncp = len(stats)
        rows = np.zeros((max(ncp, 1), 16), dtype=np.int64)
        L.asuka_jk_last_counts(h, rows.ctypes.data_as(_lib._i64p), ncp, sptr)
        class_rows = [(s_["class"], rows[i, :8].tolist(), int(rows[i, 8])) for i, s_ in enumerate(stats)]
        work = metrics.build_work(class_rows)
        f32 = sum(metrics.flop_eri(*s_["class"], int(rows[i, 10])) for i, s_ in enumerate(stats))
        q32 = int(rows[:, 9].sum())
        torch.cuda.synchronize(dev)
        kern_ms = []
        t0 = time.perf_counter()
        for _ in range(args.steps):
            build(_lib.TIME)
            ms = ctypes.c_double()
            nl = ctypes.c_int64()
            _lib.check(L.asuka_jk_last_timing(h, ctypes.byref(ms), ctypes.byref(nl)))
            kern_ms.append(ms.value)
        torch.cuda.synchronize(dev)
        wall = time.perf_counter() - t0
        if fock:  # J and K of the same mode for the accuracy stamp (untimed)
            _lib.check(L.asuka_jk_build(h, _lib.ptr(D), _lib.ptr(J), _lib.ptr(K), flags, sptr))
            torch.cuda.synchronize(dev)
        r = {"J": J, "K": K, "F": F if fock else None, "work": work, "f32": f32, "q32": q32,
             "s_per_build": wall / args.steps, "kern_s": float(np.mean(kern_ms)) / 1e3, "nlaunch": int(nl.value)}
        log(f"{mode}: {r['s_per_build'] * 1e3:.2f} ms/build, ERI kernels {r['kern_s'] * 1e3:.2f} ms over "
            f"{r['nlaunch']} launches, quartets={work.quartets} (FP32 {q32}), EFM-v1 GFLOP={work.flops / 1e9:.2f}")
        return r

    modes = ["fp64", "mixed"] if args.mode == "both" else [args.mode]
    res = {m: run_mode(m) for m in modes}
    main_mode = "mixed" if "mixed" in res else "fp64"
    r = res[main_mode]
    work = r["work"]
    s_per_build = r["s_per_build"]
    kern_s = r["kern_s"]
    value = work.quartets / s_per_build
    f32 = r["f32"]
    f64 = work.flops - f32
    achieved = work.flops / kern_s / 1e12
    accuracy = None
    if "fp64" in res and "mixed" in res:
        dJ = res["mixed"]["J"] - res["fp64"]["J"]
        dK = res["mixed"]["K"] - res["fp64"]["K"]
        if fock:  # E = 1/2 sum D (h + F) at fixed D
            dF = res["mixed"]["F"] - res["fp64"]["F"]
            dE = float(0.5 * (D * dF).sum())
        else:  # E_2e of the J/K build's density (D = C C^T, closed-shell form)
            dE = float(0.5 * (D * dJ).sum() - 0.25 * (D * dK).sum())
        accuracy = {"max_abs_dJ": float(dJ.abs().max()), "max_abs_dK": float(dK.abs().max()),
                    "max_abs_dF": float(dF.abs().max()) if fock else None,
                    "dE_Eh": dE, "tau_mixed": args.tau,
                    "fp64_s_per_build": res["fp64"]["s_per_build"],
                    "mixed_s_per_build": res["mixed"]["s_per_build"],
                    "speedup_mixed_vs_fp64": res["fp64"]["s_per_build"] / res["mixed"]["s_per_build"],
                    "fp32_quartet_fraction": r["q32"] / max(1, work.quartets)}
        log("mixed vs fp64:", json.dumps(accuracy))
    nl = type("N", (), {"value": r["nlaunch"]})()

    # parity stamp: sampled tiles vs the oracle
    parity = {}
    cpu = None
    try:
        from oracle import oracle as O

        Q = ctx.Q_sp
        octx = O.make_context(ao)
        Dsp = O.shell_pair_dmax(octx, D_h)
        rng = np.random.default_rng(7)
        sab, scd = sample_screened(Q, Dsp, cfg["eps"], ctx.eps_density, 400, rng)
        tiles = eri_tiles(ctx, sab, scd)
        dmax = 0.0
        for t, a, c in zip(tiles, sab, scd):
            dmax = max(dmax, float(np.max(np.abs(t - O.eri_tile(octx, int(a), int(c), hp=True)))))
        parity = {"sampled_tiles": int(sab.size), "max_abs_dtile_vs_oracle": dmax}
        if not args.no_cpu:
            threads = len(os.sched_getaffinity(0))
            # calibrate on a small sample (after a warm-up call: thread pool,
            # page faults), then size the sample to ~cpu_seconds
            cab, ccd = sample_screened(Q, Dsp, cfg["eps"], ctx.eps_density, 2000, rng)
            O.direct_jk_tasks(octx, D_h, cab, ccd, hp=False, threads=threads)
            cab, ccd = sample_screened(Q, Dsp, cfg["eps"], ctx.eps_density, 100000, rng)
            t1 = time.perf_counter()
            O.direct_jk_tasks(octx, D_h, cab, ccd, hp=False, threads=threads)
            rate0 = cab.size / (time.perf_counter() - t1)
            nsamp = int(min(ctx.ntasks, max(2000, rate0 * args.cpu_seconds)))
            cab, ccd = sample_screened(Q, Dsp, cfg["eps"], ctx.eps_density, nsamp, rng)
            t1 = time.perf_counter()
            O.direct_jk_tasks(octx, D_h, cab, ccd, hp=False, threads=threads)
            tc = time.perf_counter() - t1
            cpu = {"value": cab.size / tc, "unit": "shell_quartets/s", "cores": threads, "kind": "port",
                   "sample": f"{cab.size} quartets drawn uniformly from the same screened set "
                             f"(oracle, reference double-precision roots), {tc:.1f} s on {cpu_model()}",
                   "s_per_build_extrapolated": work.quartets / (cab.size / tc)}
    except Exception as e:  # the baseline is reported, never required
        log("oracle leg skipped:", repr(e))

    out = {
        "metric": METRIC,
        "value": value,
        "unit": "shell_quartets/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": s_per_build * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64" if main_mode == "fp64" else "f64+f32",
        "data": "synthetic (seeded water cluster geometry, random orthonormal occupied orbitals)",
        "config": {"workload": cfg["workload"], "nao": nao, "basis": cfg["basis"], "n_water": cfg["n"],
                   "geometry_seed": cfg["seed"], "density_seed": cfg["cseed"], "eps_schwarz": cfg["eps"],
                   "eps_density": ctx.eps_density, "mode": main_mode,
                   "launch_shapes": "autotuned" if t_tune is not None else "default"},
        "s_per_build": s_per_build,
        "quartets_per_build": work.quartets,
        "schwarz_screened_quartets": ctx.ntasks,
        "prim_quartet_roots_per_s": work.prim_roots / s_per_build,
        "context_build_s": t_ctx,
        "autotune_s": t_tune,
        "mixed_vs_fp64": accuracy,
        "roofline": {"bound": "valu", "achieved": achieved, "peak": metrics.PEAK_FP64_TFLOPS, "unit": "TFLOP/s",
                     "frac": metrics.roofline_fraction(f64, f32, kern_s),
                     "traffic": HBM_TRAFFIC.get((args.config, main_mode)),
                     "flops_fp64": f64, "flops_fp32": f32,
                     "flops_per_build_efm_v1": work.flops, "kernel_ms_per_build": kern_s * 1e3,
                     "kernel_launches_per_build": int(nl.value),
                     "alg_bytes_per_build": work.bytes_alg},
        "parity": parity,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
