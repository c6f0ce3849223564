"""Benchmark of the perturbation-stage hot path on one MI355X.

A step is one pass of the hot path over one batch: every (parameter set, k) instance of the
workload integrated with Rodas5P from τ_ini(k) to τ0 in one persistent launch, followed by the
P(k) reduction.  The default workload ("lhs128") is the batched uneven solve: 128 Latin-hypercube
parameter sets (seed 0) over h, ω_b, ω_c, ln 10¹⁰A_s, n_s and τ_reion (met by a root-find on
z_re) of ΛCDM with massless ν and lmax 10 (n = 38), × 512 k log-spaced in [1e-5, 10]/Mpc =
65,536 instances, rtol = atol = 1e-5.  The others: "massive_nu100" (n = 100, 1024 k in
[1e-5, 10]/Mpc, CMB sources on a 2000-point cosgrid of τ), "massless10" (n = 38, 100 k in
[1e-4, 1]/Mpc), "cmb" (C_l pipeline) and "bg" (batched background solve).  Inputs (background
tables, instances) are resident in device memory before the timed region; outputs stay there.

Prints ONE JSON line with the metric, the roofline of the dominant kernel (algorithmic FP64
flops counted by the emitter × the solver's own step counts, over the kernel's HIP-event time,
against the FP64 vector peak) and a CPU baseline from the oracle on a bounded sample of the same
workload on all host cores.
"""
import argparse
import json
import math
import os
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT / "symboltz.jl_amd"))
sys.path.insert(0, str(ROOT))

FP64_PEAK_TFLOPS = 78.6
HBM_PEAK_GBS = 8000.0
BOUND = "fp64-valu"   # FP64 vector ALU: the ODE kernels are neither HBM- nor MFMA-shaped


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def measured_traffic(workload, kname, n_tau):
    """HBM bytes per launch of the dominant kernel from the committed rocprofv3 FETCH_SIZE/WRITE_SIZE
    summary of the same workload and kernel (scaled by a copy-kernel calibration there), or None."""
    f = ROOT / "profiles" / f"{kname}_{workload}_hbm.md"
    if not f.exists() or (workload == "massive_nu100" and n_tau != 2000):
        return None
    import re
    m = re.search(r"traffic_bytes_per_launch:\s*(\d+)", f.read_text())
    return (int(m.group(1)), str(f.relative_to(ROOT))) if m else None


def host_cpu():
    """(threads this process may use, CPU model name from lscpu)."""
    try:
        n = len(os.sched_getaffinity(0))
    except AttributeError:
        n = os.cpu_count() or 1
    model = ""
    try:
        import subprocess
        out = subprocess.run(["lscpu"], capture_output=True, text=True, timeout=10).stdout
        for ln in out.splitlines():
            if ln.startswith("Model name:"):
                model = ln.split(":", 1)[1].strip()
    except Exception:
        pass
    return n, model


def bench_bg(args):
    """Background workload: solvebg (src/solve.jl:427-435) for a Latin-hypercube batch of cosmologies
    of the surrogate model (ΛCDM + 3 degenerate massive ν, lmax 6; 7 parameters incl. z_re and Σm_ν),
    all sets in one GPU launch, one thread per set.  A step is one batched solve."""
    from symboltz import _lib
    from symboltz.surrogate import lhs, pars_of, surrogate_model
    from symboltz.models import make_params
    M = surrogate_model()
    X = lhs(args.bg_sets, seed=1)
    Ps = np.stack([make_params(M, pars_of(M, x)) for x in X])
    for w in range(args.warmup):
        _lib.background_solve_batch(Ps)
    t_start = time.perf_counter()
    kms = []
    for s in range(args.steps):
        P, knots, nrej, status, ms = _lib.background_solve_batch(Ps)
        kms.append(ms)
        log(f"step {s}: kernel {ms:.1f} ms, {int(np.sum(status != 0))} failed")
    elapsed = (time.perf_counter() - t_start) / args.steps
    ok = bool(np.all(status == 0))
    cpu = None
    if not args.no_cpu:
        from concurrent.futures import ThreadPoolExecutor
        sel = np.arange(0, len(Ps), max(1, len(Ps) // 512))
        cores = host_cpu()[0]
        t_c = time.perf_counter()
        with ThreadPoolExecutor(max_workers=cores) as ex:
            host = list(ex.map(lambda i: _lib.background_solve(Ps[i]), sel))
        tc = time.perf_counter() - t_c
        dt0 = max(abs(P[i, 23] / h[0][23] - 1) for i, h in zip(sel, host))
        cpu = {"value": len(sel) / tc, "unit": "background solves/s", "cores": cores, "kind": "port",
               "sample": f"{len(sel)} of the sets with the host build of the same solver on {cores} threads "
                         f"({tc:.2f} s); max |tau0_gpu/tau0_cpu - 1| = {dt0:.1e}"}
        log(f"cpu baseline {len(sel) / tc:.1f} solves/s; tau0 parity {dt0:.1e}")
    kern = float(np.mean(kms))
    out = {
        "metric": "background solves/s (solvebg: Rodas5P on a, X_H+, X_He+, dT, kappa with the a = 1 event)",
        "value": len(Ps) / elapsed, "unit": "background solves/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": elapsed * 1e3, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f64", "data": "synthetic",
        "config": {"workload": "bg", "description": f"{len(Ps)} Latin-hypercube sets (seed 1) of LCDM + 3 massive nu, "
                   "rtol = atol = 1e-7, one thread per set", "n_sets": len(Ps), "all_success": ok,
                   "kernel_ms": kern, "knots_mean": float(np.mean([len(k[0]) for k in knots]))},
        "roofline": None,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)


def bench_mixed(args):
    """Mixed-model stream: requests spread over five model families (n = 38, 54, 56, 100, 236;
    symboltz.workloads.mixed_requests), every request solved: per step, the background solves of all
    requests, one plan per family and one persistent launch over all families (sbz_multi_kernel, or
    sbz_multi_kernel_rows when large plans run in several-modes bodies),
    then P(k) on 256 k in [1e-4, 5]/Mpc.  The surrogate, which answers the in-box requests of its
    family in the full stream, is measured on its own (spectra/s of inference); here every request
    takes the solver path, the stream's fallback.  The same requests are also run with one launch
    per family, for the launch-tail comparison."""
    from symboltz import surrogate as sg, workloads
    from symboltz.constants import k_dimensionless
    from symboltz.stream import MixedStream
    fams = workloads.mixed_families()
    ks = sg.k_grid()
    reqs = workloads.mixed_requests(args.mixed_sets, seed=0)
    st = MixedStream(fams, ks)
    for w in range(args.warmup):
        _, _, ret, tm = st.solve(reqs)
        log(f"warmup {w}: one launch {tm['solve_ms']:.1f} ms")
    tms = []
    t_start = time.perf_counter()
    for s in range(args.steps):
        lp, _, ret, tm = st.solve(reqs)
        tms.append(tm)
        log(f"step {s}: prepare {tm['prepare_ms']:.1f} ms, one launch {tm['solve_ms']:.1f} ms, P(k) {tm['pk_ms']:.1f} ms")
    elapsed = (time.perf_counter() - t_start) / args.steps
    _, _, ret2, tm2 = st.solve(reqs, mode="serial")
    _, _, ret3, tm3 = st.solve(reqs, mode="concurrent")
    kern_ms = float(np.mean([t["solve_ms"] for t in tms]))
    # the launch takes the several-modes kernel when a family with such a body has a plan of at
    # least twice the capacity (sbz_multi_run)
    cap, rows = st.multi.capacity(), st.multi.rows()
    per_fam = np.bincount([r[0] for r in reqs], minlength=len(fams)) * len(ks)
    kname = ("sbz_multi_kernel_rows" if any(rw > 1 and m >= 2 * cap for rw, m in zip(rows, per_fam))
             else "sbz_multi_kernel")
    flops = tms[-1]["flops"]
    achieved = flops / (kern_ms * 1e-3) / 1e12
    ok = bool(np.all(ret == 0))
    n_modes = len(reqs) * len(ks)
    log(f"one launch {kern_ms:.1f} ms vs one launch per family {tm2['solve_ms']:.1f} ms (serial), "
        f"{tm3['solve_ms']:.1f} ms (concurrent streams); all success {ok}; "
        f"{achieved:.3f} TFLOP/s")
    # error against the same solver at rtol = atol = 1e-8, on the first two requests of every family
    ref_err = None
    if not args.no_ref:
        sub = [i for f in range(len(fams)) for i in [j for j, r in enumerate(reqs) if r[0] == f][:2]]
        ref = MixedStream(fams, ks, rtol=1e-8, atol=1e-8)
        try:
            lpr, _, retr, _ = ref.solve([reqs[i] for i in sub])
        finally:
            ref.close()
        err = np.abs(np.expm1(lp[sub] - lpr)).max(axis=1)
        ref_err = {f"n={fams[reqs[i][0]].n_unknowns}": 0.0 for i in sub}
        for e, i in zip(err, sub):
            key = f"n={fams[reqs[i][0]].n_unknowns}"
            ref_err[key] = max(ref_err[key], float(e))
        log(f"max |P(rtol 1e-5) / P(rtol 1e-8) - 1| by family (first two requests each): {ref_err} "
            f"(reference run success: {bool(np.all(retr == 0))})")
    cpu = None
    if not args.no_cpu:
        from oracle import oracle
        from symboltz.problem import CosmologyProblem, solvebg
        cores, cpu_model = host_cpu()
        tc, par, n_cpu = 0.0, 0.0, 0
        for f in range(len(fams)):   # the first request of every family, all its k
            i = next(j for j, r in enumerate(reqs) if r[0] == f)
            M, p = fams[f], reqs[i][1]
            b = solvebg(CosmologyProblem(M, p))
            O = oracle.Oracle(M, b.P, b.t, b.u, b.du)
            k = k_dimensionless(ks, p["h"])
            t0 = np.clip(np.minimum(1e-2 / k, 1e-4), b.t[0], b.t[-1])
            t_c = time.perf_counter()
            o = O.solve(k, t0, nthreads=cores, sparse=True)
            tc += time.perf_counter() - t_c
            po = np.array([O.p0(kk) * O.delta_m(kk, b.t[-1], o["u"][j]) ** 2 for j, kk in enumerate(k)])
            par = max(par, float(np.max(np.abs(np.exp(lp[i]) / po - 1))))
            n_cpu += 1
        cpu = {"value": n_cpu / tc, "unit": "spectra/s", "cores": cores, "kind": "port", "solver": "sparse",
               "cpu": cpu_model, "nproc": os.cpu_count(),
               "sample": f"the first request of each of the {len(fams)} families ({n_cpu} spectra x {len(ks)} k), "
                         f"{tc:.1f} s on {cores} host threads ({cpu_model}); oracle sparse path; "
                         f"max |P_gpu/P_cpu - 1| = {par:.2e}"}
        log(f"cpu baseline: {cpu['value']:.3f} spectra/s, parity {par:.2e}")
    out = {
        "metric": "spectra/s of the mixed-model stream (every request solved, P(k) on 256 k)",
        "value": len(reqs) / elapsed, "unit": "spectra/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": elapsed * 1e3, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f64", "data": "synthetic",
        "config": {"workload": "mixed",
                   "description": f"{len(reqs)} requests over {len(fams)} model families "
                                  f"(n = {[M.n_unknowns for M in fams]}), Latin-hypercube parameters (seed 0), "
                                  f"256 k log-spaced in [1e-4, 5]/Mpc: {n_modes} k-modes in one launch; rtol=atol=1e-5",
                   "k_mode_solves_per_s": n_modes / elapsed, "all_success": ok,
                   "pk_max_rel_err_vs_rtol_1e-8": None if ref_err is None else max(ref_err.values()),
                   "pk_rel_err_vs_rtol_1e-8_by_family": ref_err,
                   "one_launch_kernel_ms": kern_ms, "per_family_launches_kernel_ms": tm2["solve_ms"],
                   "per_family_concurrent_ms": tm3["solve_ms"],
                   "stage_ms_per_step": {k: float(np.mean([t[k] for t in tms])) for k in
                                         ("prepare_ms", "solve_ms", "pk_ms")}},
        "roofline": {"bound": BOUND, "achieved": achieved, "peak": FP64_PEAK_TFLOPS, "unit": "TFLOP/s",
                     "frac": achieved / FP64_PEAK_TFLOPS, "traffic": None, "kernel": kname,
                     "kernel_ms": kern_ms, "flops_per_launch": flops, "steps_per_launch": tms[-1]["steps"]},
        "cpu_baseline": cpu,
    }
    st.close()
    print(json.dumps(out), flush=True)


def bench_mlp_train(args):
    """Surrogate training step: Adam steps of the 7 -> 512^4 -> 256 GELU network (BF16 MFMA products,
    FP32 accumulation and master weights) on batches of 1024 rows drawn from 32,768 resident
    synthetic rows (uniform inputs, normal targets).  A step is one epoch of 32 Adam steps."""
    from symboltz.mlp import Mlp
    widths = [7, 512, 512, 512, 512, 256]
    B, n_rows = 1024, 32768
    rng = np.random.default_rng(0)
    X = rng.uniform(-1, 1, (n_rows, widths[0])).astype(np.float32)
    T = rng.normal(0, 1, (n_rows, widths[-1])).astype(np.float32)
    net = Mlp(widths, max_batch=B, seed=0)
    net.set_data(X, T)
    for w in range(args.warmup):
        net.train(rng.permutation(n_rows), B, 1e-3)
    ms = 0.0
    t_start = time.perf_counter()
    for s in range(args.steps):
        loss, t = net.train(rng.permutation(n_rows), B, 1e-3)
        ms += t
        log(f"step {s}: {t:.2f} ms for {n_rows // B} Adam steps, loss {float(loss.mean()):.4e}")
    elapsed = (time.perf_counter() - t_start) / args.steps
    adam_steps = n_rows // B
    macs = sum(widths[l] * widths[l + 1] for l in range(len(widths) - 1))
    flops = 6.0 * B * macs * adam_steps        # forward, backward data and backward weight products
    achieved = flops / (ms / args.steps * 1e-3) / 1e12
    cpu = None
    if not args.no_cpu:
        from oracle import mlp_oracle as mo
        p = net.get_params()
        t_c = time.perf_counter()
        mo.loss_and_grad(p, widths, X[:B], T[:B])
        tc = time.perf_counter() - t_c
        cores, cpu_model = host_cpu()
        cpu = {"value": 1.0 / tc, "unit": "Adam steps/s (gradient only)", "cores": cores, "kind": "port",
               "sample": f"one gradient of a 1024-row batch with the numpy oracle (FP64 BLAS on up to {cores} "
                         f"threads, BF16 rounding points; {cpu_model}), {tc:.2f} s"}
    net.close()
    out = {
        "metric": "surrogate training Adam steps/s (batch 1024, 7 -> 512x4 -> 256)",
        "value": adam_steps / elapsed, "unit": "Adam steps/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": elapsed * 1e3, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "bf16", "data": "synthetic",
        "config": {"workload": "mlp_train", "description": f"{n_rows} resident rows, batch {B}, {adam_steps} Adam "
                   f"steps per step, widths {widths}", "gpu_ms_per_adam_step": ms / args.steps / adam_steps},
        "roofline": {"bound": "mfma", "achieved": achieved, "peak": 2500.0, "unit": "TFLOP/s",
                     "frac": achieved / 2500.0, "traffic": None, "kernel": "training step (all kernels)",
                     "flops_per_adam_step": 6.0 * B * macs},
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)


def bench_cmb(args):
    """C_l workload: spectrum_cmb([TT, EE, ψψ]) of the reference's CLASS-comparison model (ΛCDM
    lmax 5; test/runtests.jl:14-16,879-886) for a batch of cosmologies: set 0 is Planck18 (checked
    against test/class_Cl.dat), the others a seeded Latin hypercube over h, ω_b, ω_c, ln 1e10 As, ns.
    D_l at every l in [2, 2500] from 152 cached l.  A step is the whole pipeline for the batch:
    background solves (host), 131 source modes per set on 300 τ in one launch, the device C_l stage
    per set (Bessel table, k interpolation to ~10,000 k, LOS, Limber, k integral) and the l spline."""
    from scipy.stats import qmc
    from symboltz import models
    from symboltz.bessel import SphericalBesselCache
    from symboltz.cmb import spectrum_cmb_sets
    from symboltz.problem import CosmologyProblem, remake, spectrum_primordial
    M = models.LCDM(lmax=5)
    pars = models.parameters_Planck18(M)
    base = CosmologyProblem(M, pars)
    X = qmc.scale(qmc.LatinHypercube(d=5, seed=0).random(max(args.cmb_sets - 1, 1)),
                  [0.60, 0.020, 0.10, 2.9, 0.92], [0.75, 0.024, 0.14, 3.2, 1.00])
    probs = [base] + [remake(base, dict(h=h, Omega_b0=wb / h**2, Omega_c0=wc / h**2, ln_As1e10=la, ns=ns))
                      for h, wb, wc, la, ns in X[:args.cmb_sets - 1]]
    d = np.loadtxt(ROOT / "tests" / "golden" / "class_Cl.dat")
    lsc = d[:, 0]
    ls = np.unique(np.round(np.exp(np.linspace(np.log(lsc[0]), np.log(lsc[-1]), 200))).astype(int))
    jl = SphericalBesselCache(ls)
    modes = ["TT", "EE", "ψψ"]
    for w in range(args.warmup):
        spectrum_cmb_sets(modes, probs, jl, lsc, normalization="Dl")
        log(f"warmup {w} done")
    tms = []
    t_start = time.perf_counter()
    for s in range(args.steps):
        tm, det = {}, []
        D = spectrum_cmb_sets(modes, probs, jl, lsc, normalization="Dl", timing=tm, details=det)
        tms.append(tm)
        log(f"step {s}: " + ", ".join(f"{k} {v:.2f}" for k, v in tm.items() if k.endswith("_ms")))
    elapsed = (time.perf_counter() - t_start) / args.steps
    nw = lambda a, b: float(np.linalg.norm(a - b) / max(np.linalg.norm(a), np.linalg.norm(b)))
    vs_class = {m: nw(D[0][:, i], d[:, 1 + i]) for i, m in enumerate(modes)}
    log(f"set 0 vs class_Cl.dat: {vs_class}")
    # LOS kernel: per (k, τ, l) it reads 4 table doubles (32 B) and does a Hermite read (12 flops)
    # plus 2 flops per source
    los_ms = float(np.mean([t["los_ms"] for t in tms])) / len(probs)
    nk, nt, nl = len(det[0]["ks_fine"]), len(det[0]["taus"]), len(ls)
    src_ms = float(np.mean([t["sources_ms"] for t in tms]))
    gm = base.gpu
    steps = int(np.sum(tms[-1]["steps"]))
    flops = steps * gm.em.flops_step
    achieved = flops / (src_ms * 1e-3) / 1e12
    cpu = None
    if not args.no_cpu:
        from oracle import oracle, oracle_cmb
        from symboltz.problem import solvebg
        bg = solvebg(base)
        O = oracle.Oracle(M, bg.P, bg.t, bg.u, bg.du)
        t_c = time.perf_counter()
        Do, _ = oracle_cmb.spectrum_cmb(O, lambda k: spectrum_primordial(k, base), modes, ls, ls_fine=lsc)
        tc = time.perf_counter() - t_c
        par = {m: nw(D[0][:, i], Do[:, i]) for i, m in enumerate(modes)}
        cpu = {"value": 1.0 / tc, "unit": "C_l spectrum sets/s", "cores": oracle_cmb._threads(), "kind": "port",
               "sample": f"set 0 (Planck18) once, {tc:.2f} s: oracle source solve, C++ interpolation and LOS, "
                         f"scipy k integral; norm-wise |D_gpu - D_cpu| for set 0 = {par}"}
        log(f"cpu baseline {tc:.2f} s per set; parity {par}")
    out = {
        "metric": "CMB C_l spectrum sets/s (TT, EE, psipsi at l = 2..2500, whole pipeline)",
        "value": len(probs) / elapsed, "unit": "C_l spectrum sets/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": elapsed * 1e3, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f64", "data": "synthetic",
        "config": {"workload": "cmb", "description": f"LCDM lmax 5; {len(probs)} cosmologies (Planck18 + LHS seed 0) x "
                   f"{len(det[0]['ks_coarse'])} Chebyshev source modes x {nt} tau in one launch; ~{nk} fine k; "
                   f"{nl} cached l; D_l at l = 2..2500",
                   "n_sets": len(probs), "set0_vs_class_Cl_normwise": vs_class,
                   "stage_ms_per_step": {k: float(np.mean([t[k] for t in tms])) for k in tms[0] if k.endswith("_ms")},
                   "los_kernel": {"kernel": "sbz_cmb_los_kernel", "ms_per_set": los_ms,
                                  "table_bytes_read_per_set": 32.0 * nk * nt * nl,
                                  "table_read_GBs": 32.0 * nk * nt * nl / (los_ms * 1e-3) / 1e9,
                                  "flops_per_set": 20.0 * nk * nt * nl,
                                  "TFLOPs": 20.0 * nk * nt * nl / (los_ms * 1e-3) / 1e12}},
        "roofline": {"bound": BOUND, "achieved": achieved, "peak": FP64_PEAK_TFLOPS, "unit": "TFLOP/s",
                     "frac": achieved / FP64_PEAK_TFLOPS, "traffic": None, "kernel": "source solve (wave kernel)",
                     "kernel_ms": src_ms, "flops_per_launch": flops, "flops_per_step": gm.em.flops_step,
                     "steps_per_launch": steps},
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)




def prepare(args, M, sets, ks_mpc, gm):
    """Background solves and uploads; returns (backgrounds, inst_set, k, τ_start, seconds per set)."""
    from symboltz.constants import k_dimensionless
    from symboltz.problem import CosmologyProblem, solvebg, solvebg_sets
    t_bg = time.perf_counter()
    bgs, inst_set, ks_l, t0_l = [], [], [], []
    if len(sets) > 1:   # all backgrounds at once on the GPU, one thread per set
        bg_all = solvebg_sets([CosmologyProblem(M, pars) for pars in sets])
    for si, pars in enumerate(sets):
        bg = bg_all[si] if len(sets) > 1 else solvebg(CosmologyProblem(M, pars))
        gm.upload_background(si, bg.t, bg.u, bg.du, bg.P)
        k_s = k_dimensionless(ks_mpc, pars["h"])
        bgs.append(bg)
        inst_set.append(np.full(len(k_s), si, dtype=np.int32))
        ks_l.append(k_s)
        t0_l.append(np.clip(np.minimum(1e-2 / k_s, 1e-4), bg.t[0], bg.t[-1]))
    t_bg = (time.perf_counter() - t_bg) / len(sets)
    return bgs, np.concatenate(inst_set), np.concatenate(ks_l), np.concatenate(t0_l), t_bg


def side_kernel_ms(args):
    """Kernel time of the n = 100 massive-ν workload with its sources (a few runs), reported beside
    the main workload for continuity."""
    from symboltz import workloads
    from symboltz.solver import GpuModel, SAVE_SOURCES, ALGO_WAVE
    M, sets, ks_mpc, _ = workloads.workload("massive_nu100")
    gm = GpuModel(M)
    bgs, iset, ks, t0, _ = prepare(args, M, sets, ks_mpc, gm)
    taus = workloads.source_taus(bgs[0].t[0], bgs[0].t[-1], 2000)
    plan = gm.plan(iset, ks, t0, saveat=taus, save=SAVE_SOURCES, algo=ALGO_WAVE)
    plan.run()
    ms = [plan.run() for _ in range(3)]
    ok = bool(np.all(plan.fetch()["retcode"] == 0))
    plan.close()
    gm.close()
    return float(np.mean(ms)), ok


def side_batched_ms(args):
    """Kernel time of the same n = 100 model over 64 parameter sets × the 1024 k (P(k) only), with
    the autotuned kernel variant: the model at a size that fills the device."""
    from symboltz import workloads
    from symboltz.autotune import NAMES
    from symboltz.solver import GpuModel, ALGO_AUTO
    M, sets, ks_mpc, _ = workloads.workload("massive_nu100_lhs64")
    gm = GpuModel(M)
    bgs, iset, ks, t0, _ = prepare(args, M, sets, ks_mpc, gm)
    plan = gm.plan(iset, ks, t0, algo=ALGO_AUTO)
    plan.run()
    ms = [plan.run() for _ in range(2)]
    ok = bool(np.all(plan.fetch()["retcode"] == 0))
    name = NAMES.get(plan.algo, str(plan.algo))
    plan.close()
    gm.close()
    return float(np.mean(ms)), ok, name, len(ks)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", default="lhs128", choices=["lhs128", "massive_nu100", "massive_nu100_lhs64", "massless10", "cmb", "bg", "mixed",
                                                      "mlp_train"])
    ap.add_argument("--lanes", type=int, default=0)
    ap.add_argument("--algo", choices=["auto", "wave", "wave2", "wave4", "wave2h", "wave_altlu"], default="auto",
                    help="auto: timed choice of the kernel variant (default); wave: one wavefront per "
                         "k-mode, one resident per SIMD; wave2: two resident per SIMD; wave4: four k-modes per "
                         "wavefront; wave2h: two k-modes per wavefront; wave_altlu: wave with the other border-factor "
                         "layout")
    ap.add_argument("--graph", action="store_true", help="replay each step from a captured HIP graph")
    ap.add_argument("--cpu-sample", type=int, default=0,
                    help="oracle baseline: every N-th instance (0: all for one set, every 4th for many sets)")
    ap.add_argument("--cpu-dense", action="store_true",
                    help="oracle baseline with its dense path (J by unit vectors, partial-pivot LU)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-side", action="store_true", help="skip the n = 100 side measurement")
    ap.add_argument("--no-ref", action="store_true", help="skip the rtol 1e-8 accuracy run")
    ap.add_argument("--no-sources", action="store_true", help="massive_nu100: P(k) only, no CMB sources")
    ap.add_argument("--n-tau", type=int, default=2000, help="points of the source time grid")
    ap.add_argument("--cmb-sets", type=int, default=16, help="cosmologies per step of the cmb workload")
    ap.add_argument("--bg-sets", type=int, default=16384, help="cosmologies per step of the bg workload")
    ap.add_argument("--mixed-sets", type=int, default=320, help="requests per step of the mixed workload")
    ap.add_argument("--pmc-copy", type=float, default=0.0,
                    help="only run a plain device copy of this many MiB (HBM counter calibration) and exit")
    args = ap.parse_args()
    if args.pmc_copy > 0:
        import torch
        n = int(args.pmc_copy * 2**20) // 8
        x = torch.ones(n, dtype=torch.float64, device="cuda")
        y = torch.empty_like(x)
        torch.cuda.synchronize()
        for _ in range(3):
            y.copy_(x)
        torch.cuda.synchronize()
        print(json.dumps({"copy_mib": args.pmc_copy, "bytes_read_per_copy": 8 * n, "bytes_written_per_copy": 8 * n}))
        return
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for {args.gpus}): instances are independent and would be "
              f"sharded across GPUs by cost, which is not measured here.", file=sys.stderr)
        raise SystemExit(2)

    from symboltz import _lib, workloads
    from symboltz.solver import (GpuModel, ALGO_WAVE, ALGO_WAVE2, ALGO_WAVE4, ALGO_WAVE2H, ALGO_WAVE_ALTLU, ALGO_AUTO,
                                 SAVE_SOURCES)

    if _lib.device_count() < 1:
        raise SystemExit("bench.py: no GPU visible")
    if args.workload == "cmb":
        return bench_cmb(args)
    if args.workload == "bg":
        return bench_bg(args)
    if args.workload == "mixed":
        return bench_mixed(args)
    if args.workload == "mlp_train":
        return bench_mlp_train(args)
    M, sets, ks_mpc, desc = workloads.workload(args.workload)
    gm = GpuModel(M)
    bgs, inst_set, ks, t0, t_bg = prepare(args, M, sets, ks_mpc, gm)
    bg = bgs[0]
    algo = {"auto": ALGO_AUTO, "wave": ALGO_WAVE, "wave2": ALGO_WAVE2, "wave4": ALGO_WAVE4, "wave2h": ALGO_WAVE2H,
            "wave_altlu": ALGO_WAVE_ALTLU}[args.algo]
    taus = None
    if not args.no_sources and args.workload == "massive_nu100":
        taus = workloads.source_taus(bg.t[0], bg.t[-1], args.n_tau)
    plan = gm.plan(inst_set, ks, t0, lanes=args.lanes, algo=algo, saveat=taus,
                   save=SAVE_SOURCES if taus is not None else 0)
    chosen = {0: "wave", ALGO_WAVE: "wave", ALGO_WAVE2: "wave2", ALGO_WAVE4: "wave4", ALGO_WAVE2H: "wave2h",
              ALGO_WAVE_ALTLU: "wave_altlu"}[plan.algo]
    kname = {"wave": "sbz_wave_kernel", "wave2": "sbz_wave_kernel_o2", "wave4": "sbz_wave4_kernel",
             "wave2h": "sbz_wave2h_kernel", "wave_altlu": "sbz_wave_kernel_altlu"}[chosen]
    from symboltz import autotune
    if args.algo == "auto" and autotune.last_times:
        log("autotune sample times (ms): " + ", ".join(f"{k} {v:.1f}" for k, v in autotune.last_times.items()))
    log(f"workload {args.workload}: n={gm.n}, {len(sets)} set(s), {len(ks)} instances, background "
        f"{t_bg*1e3:.1f} ms per set ({len(bg.t)} knots in set 0); kernel variant {chosen}"
        f"{' (graph replays)' if args.graph else ''}")
    run = plan.run_graph if args.graph else plan.run
    for w in range(args.warmup):
        ms = run()
        plan.pk()
        log(f"warmup {w}: {'graph' if args.graph else 'kernel'} {ms:.1f} ms")
    kernel_ms = []
    t_start = time.perf_counter()
    for s in range(args.steps):
        kernel_ms.append(run())
        pk = plan.pk()
        log(f"step {s}: {'graph' if args.graph else 'kernel'} {kernel_ms[-1]:.1f} ms")
    elapsed = time.perf_counter() - t_start
    res = plan.fetch()
    src = plan.sources() if taus is not None else None
    ms_per_step = elapsed / args.steps * 1e3
    n_modes = len(ks)
    value = n_modes / (elapsed / args.steps)
    steps_total = int(res["stats"][:, 0].sum() + res["stats"][:, 1].sum())
    flops = steps_total * gm.em.flops_step
    kern_s = float(np.mean(kernel_ms)) * 1e-3
    achieved = flops / kern_s / 1e12
    ok = bool(np.all(res["retcode"] == 0))
    log(f"retcodes ok={ok}; steps/mode min {res['stats'][:,0].min()} max {res['stats'][:,0].max()} "
        f"total {steps_total}; flops/step {gm.em.flops_step}; {achieved:.3f} TFLOP/s")

    # accuracy of the rtol = 1e-5 spectrum against the same GPU solver at rtol = atol = 1e-8 (outside
    # the timed region; high-k modes need ~15x the steps of rtol 1e-5, beyond the default maxiters)
    pk_err = None
    if not args.no_ref:
        ref = gm.plan(inst_set, ks, t0, rtol=1e-8, atol=1e-8, maxiters=2_000_000, lanes=args.lanes, algo=plan.algo)
        ref.run()
        pk_ref = ref.pk()
        ref_ok = bool(np.all(ref.fetch()["retcode"] == 0))
        ref.close()
        rel = np.abs(pk / pk_ref - 1)
        pk_err = float(np.max(rel))
        kmpc = np.concatenate([ks_mpc] * len(sets))
        bands = {}
        for lo_, hi_ in ((1e-5, 1e-2), (1e-2, 0.1), (0.1, 1.0), (1.0, 10.0 + 1e-9)):
            m = (kmpc >= lo_) & (kmpc < hi_)
            if m.any():
                bands[f"[{lo_:g},{min(hi_, 10):g}]/Mpc"] = float(rel[m].max())
        log(f"max |P(rtol 1e-5) / P(rtol 1e-8) - 1| = {pk_err:.2e} (reference run success: {ref_ok}); by k band {bands}")

    side = None
    if not args.no_side and args.workload == "lhs128":
        side_ms, side_ok = side_kernel_ms(args)
        side = {"workload": "massive_nu100 (n=100, 1024 k, sources on 2000 tau)", "kernel": "sbz_wave_kernel",
                "kernel_ms": side_ms, "all_success": side_ok, "k_mode_solves_per_s_kernel": 1024 / (side_ms * 1e-3)}
        log(f"side: massive_nu100 kernel {side_ms:.1f} ms")
        b_ms, b_ok, b_algo, b_n = side_batched_ms(args)
        side["batched"] = {"workload": "massive_nu100_lhs64 (the n=100 model, 64 sets x 1024 k, P(k) only)",
                           "algo": b_algo, "kernel_ms": b_ms, "all_success": b_ok,
                           "k_mode_solves_per_s_kernel": b_n / (b_ms * 1e-3)}
        log(f"side: massive_nu100_lhs64 kernel {b_ms:.1f} ms ({b_algo}), {b_n / (b_ms * 1e-3):.0f} k-mode solves/s")

    cpu = None
    if not args.no_cpu:
        from oracle import oracle
        step = args.cpu_sample if args.cpu_sample > 0 else (4 if len(sets) > 1 else 1)
        sparse = not args.cpu_dense
        sel = np.arange(0, n_modes, step)
        cores, cpu_model = host_cpu()
        tc = 0.0
        po = np.zeros(len(sel))
        o_src = []
        for si in range(len(sets)):   # one oracle per parameter set, each on its sampled instances
            m = sel[inst_set[sel] == si]
            if len(m) == 0:
                continue
            b = bgs[si]
            O = oracle.Oracle(M, b.P, b.t, b.u, b.du)
            t_c = time.perf_counter()
            o = O.solve(ks[m], t0[m], nthreads=cores, saveat=taus, sources=taus is not None, sparse=sparse)
            tc += time.perf_counter() - t_c
            po[np.searchsorted(sel, m)] = [O.p0(k) * O.delta_m(k, b.t[-1], o["u"][i]) ** 2 for i, k in enumerate(ks[m])]
            if taus is not None:
                o_src.append(o["src"])
        parity = float(np.max(np.abs(pk[sel] / po - 1)))
        sp = ""
        if taus is not None:   # per mode and source: max |S_gpu - S_cpu| / max |S_cpu| over the grid
            refs = np.concatenate(o_src)
            sc = np.abs(refs).max(axis=1, keepdims=True) + 1e-300
            e = (np.abs(src[sel] - refs) / sc).max(axis=(1, 2))
            kc = 3000.0   # H0/c (≈0.67/Mpc, l ≲ k τ0 ≈ 9500)
            lo = ks[sel] <= kc
            sp = (f"; sources max |ΔS|/max|S| = {e[lo].max():.2e} over the {int(lo.sum())} modes with k <= {kc:g} H0/c "
                  f"(above it the late-time ISW term is tolerance-limited in both paths)")
        how = ("sparse path: J by column coloring of the linear RHS (one background evaluation per step), "
               "static-pattern LU in a minimum-degree order" if sparse else
               "dense path: J by unit vectors, partial-pivot LU")
        cpu = {"value": len(sel) / tc, "unit": "k-mode solves/s", "cores": cores, "kind": "port",
               "solver": "sparse" if sparse else "dense", "cpu": cpu_model, "nproc": os.cpu_count(),
               "sample": f"{'all' if step == 1 else f'every {step}th'} instance of the workload ({len(sel)} modes, "
                         f"{'with' if taus is not None else 'without'} the sources), {tc:.1f} s on {cores} host "
                         f"threads ({cpu_model}); oracle {how}, Rodas5P with the GPU path's controller; "
                         f"max |P_gpu/P_cpu - 1| = {parity:.2e}{sp}"}
        log(f"cpu baseline: {cpu['value']:.1f} solves/s, parity {parity:.2e}{sp}")

    out = {
        "metric": "k-mode solves/s at rtol 1e-5 (plain GPU Rodas5P, P(k) at z=0"
                  + ("" if taus is None else f" + CMB sources on {len(taus)} τ") + ")",
        "value": value,
        "unit": "k-mode solves/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": args.workload, "description": desc, "algo": chosen, "graph": bool(args.graph),
                   "autotune_sample_ms": dict(autotune.last_times) if args.algo == "auto" else None,
                   "n_sets": len(sets), "n_tau_sources": 0 if taus is None else len(taus), "n_modes": n_modes,
                   "n_unknowns": gm.n, "background_ms_per_set": t_bg * 1e3, "all_success": ok,
                   "pk_max_rel_err_vs_rtol_1e-8": pk_err,
                   "pk_rel_err_vs_rtol_1e-8_by_k_band": bands if pk_err is not None else None,
                   "side_measurement": side},
        "roofline": {"bound": BOUND, "achieved": achieved, "peak": FP64_PEAK_TFLOPS, "unit": "TFLOP/s",
                     "frac": achieved / FP64_PEAK_TFLOPS, "traffic": None,
                     "kernel": kname, "kernel_ms": float(np.mean(kernel_ms)),
                     "flops_per_launch": flops, "flops_per_step": gm.em.flops_step, "steps_per_launch": steps_total},
        "cpu_baseline": cpu,
    }
    tr = measured_traffic(args.workload, kname, args.n_tau)
    if tr is not None:
        out["roofline"]["traffic"] = tr[0]
        out["roofline"]["traffic_source"] = tr[1]
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
