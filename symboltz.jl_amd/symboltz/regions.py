"""Where a step of the wave kernels goes: region timers on the GPU.

    python -m symboltz.regions wave4 [--sets 32]      # sbz_wave4_kernel, the n = 38 Latin-hypercube batch
    python -m symboltz.regions wave                   # sbz_wave_kernel, the n = 100 model with sources

builds a copy of the model's code object whose kernel body accumulates `clock64()` (the shader
clock) between fixed points of the step in per-wavefront registers, runs the workload through it,
and prints each region's share of the time and its cycles per step attempt.  The copy writes the
totals over the final states of its instances (slot j = region j, last slot = -attempts), so its
results are timings, not solutions; it is cached as _kcache/regions_<kernel>.hsaco and never used
by the solver.  The clock reads are scalar and the compiler may move vector work across them, so
adjacent regions can trade a few percent; the totals are the kernel's.
"""
import argparse

import numpy as np

from .build import KCACHE, compile_code_object
from .solver import model_source

NR = 16


def _tick(j):
    return f"{{ long long tn = clock64(); PR[{j}] += tn - tq; tq = tn; }}"


def _patch(body, pairs):
    for old, new in pairs:
        if body.count(old) != 1:
            raise RuntimeError(f"region anchor not found once in the kernel source: {old[:60]!r}")
        body = body.replace(old, new)
    return body


WAVE4_REGIONS = ["pull/start", "row checks", "background", "chain factor", "W assembly", "border LU",
                 "unpack+tri2+Woodbury", "dfdt apply", "stage input", "stage apply", "stage rhs", "chain fwd",
                 "border solve", "chain bwd", "post-stage", "controller"]
WAVE_REGIONS = ["saves", "step checks", "background", "chain factor", "W assembly", "LU+Woodbury", "dfdt apply",
                "stage input", "stage apply", "stage rhs", "chain fwd", "border solve", "chain bwd", "error/exit",
                "controller"]


def wave4_source(src):
    """The model source with sbz_waverow_run (the four-mode kernel's body) instrumented; totals per
    wavefront in row 0's last instance."""
    i0 = src.index("SBZ_DEV void sbz_waverow_run")
    i1 = src.index("\ntemplate <bool FUSED, int LN>\nSBZ_DEV void sbz_waverow_body", i0)
    T = _tick
    body = _patch(src[i0:i1], [
        ("    auto finish = [&]() {", f"    long long PR[{NR}] = {{}}; long long PI = 0, tq = 0;\n    auto finish = [&]() {{"),
        ("    for (;;) {\n        // ---- rows without", "    for (;;) {\n        PI++; tq = clock64();\n        // ---- rows without"),
        ("        if (!__any(have)) break;\n", "        if (!__any(have)) break;\n        " + T(0) + "\n"),
        ("        // background at the six distinct stage times, lanes 0..5 of every row\n",
         "        " + T(1) + "\n        // background at the six distinct stage times, lanes 0..5 of every row\n"),
        ("        __syncthreads();\n        const double* B0 = L.B[0];", "        __syncthreads();\n        " + T(2) + "\n        const double* B0 = L.B[0];"),
        ("        const double xc = (c.isB && c.bchain >= 0) ? L.C[c.bchain] : 0.0;\n",
         "        " + T(3) + "\n        const double xc = (c.isB && c.bchain >= 0) ? L.C[c.bchain] : 0.0;\n"),
        ("        bool ok = sbz_wb_lu<BC>(c.lane, row, rpo);\n",
         "        " + T(4) + "\n        bool ok = sbz_wb_lu<BC>(c.lane, row, rpo);\n        " + T(5) + "\n"),
        ("        sbz_wave_apply<LOADS_FIRST>(c, B0, L.dB, L.Cs, k, Un, DT);  // ∂f/∂τ at (t, u_n)\n",
         "        " + T(6) + "\n        sbz_wave_apply<LOADS_FIRST>(c, B0, L.dB, L.Cs, k, Un, DT);  // ∂f/∂τ at (t, u_n)\n"),
        ("        if (live) st_j++;\n        __syncthreads();\n", "        if (live) st_j++;\n        __syncthreads();\n        " + T(7) + "\n"),
        ("                __syncthreads();\n                sbz_wave_apply<LOADS_FIRST>(c, L.B[sbz_stage_bg[s]], nullptr, L.Cs, k, in, Ks);\n",
         "                __syncthreads();\n                " + T(8) + "\n                sbz_wave_apply<LOADS_FIRST>(c, L.B[sbz_stage_bg[s]], nullptr, L.Cs, k, in, Ks);\n"),
        ("                if (live && (s <= 8 || EEst <= 1.0)) st_f++;\n                __syncthreads();\n",
         "                if (live && (s <= 8 || EEst <= 1.0)) st_f++;\n                __syncthreads();\n                " + T(9) + "\n"),
        ("            __syncthreads();\n            // solve W x = Ks:", "            __syncthreads();\n            " + T(10) + "\n            // solve W x = Ks:"),
        ("                L.C[c.lane] = y;\n            }\n            __syncthreads();\n", "                L.C[c.lane] = y;\n            }\n            __syncthreads();\n            " + T(11) + "\n"),
        ("                    L.X[c.lane] = x;\n                }\n            }\n            __syncthreads();\n",
         "                    L.X[c.lane] = x;\n                }\n            }\n            __syncthreads();\n            " + T(12) + "\n"),
        ("                SBZ_CHAIN_GROUPS(SBZ_CG_BWD)\n#undef SBZ_CG_BWD\n            }\n            __syncthreads();\n",
         "                SBZ_CHAIN_GROUPS(SBZ_CG_BWD)\n#undef SBZ_CG_BWD\n            }\n            __syncthreads();\n            " + T(13) + "\n"),
        ("        __syncthreads();\n        // the controller's two powers", "        __syncthreads();\n        " + T(14) + "\n        // the controller's two powers"),
    ])
    tail = "        __syncthreads();\n    }\n}\n"
    if not body.endswith(tail):
        raise RuntimeError("region anchor: end of the four-mode body")
    body = body[:-len(tail)] + ("        __syncthreads();\n        " + T(15) + "\n    }\n"
                                "    if (threadIdx.x == 0) { double* o = a.out_u + (size_t)inst * SBZ_N;"
                                f" for (int j = 0; j < {NR}; ++j) o[j] = (double)PR[j]; o[SBZ_N - 1] = -(double)PI; }}\n}}\n")
    return src[:i0] + body + src[i1:]


def wave_source(src):
    """The model source with sbz_wave_mode instrumented; totals per instance."""
    i0 = src.index("SBZ_DEV void sbz_wave_mode(")
    i1 = src.index("SBZ_DEV void sbz_wave_body", i0)
    i1 = src.rindex("\ntemplate <", i0, i1)   # the template line of sbz_wave_body
    T = _tick
    body = _patch(src[i0:i1], [
        ("        int bidx = kidx;   // knot cursor", f"        long long PR[{NR}] = {{}}; long long tq = clock64();\n        int bidx = kidx;   // knot cursor"),
        ("            if (!(ret == SBZ_RET_SUCCESS && t < t1)) break;\n", "            " + T(0) + "\n            if (!(ret == SBZ_RET_SUCCESS && t < t1)) break;\n"),
        ("            // background at the six distinct stage times, one lane each (dual: d/dτ from lane 0)\n",
         "            " + T(1) + "\n            // background at the six distinct stage times, one lane each (dual: d/dτ from lane 0)\n"),
        ("            const double* B0 = L.B[0];\n", "            " + T(2) + "\n            const double* B0 = L.B[0];\n"),
        ("            const double xc = (c.isB && c.bchain >= 0) ? L.C[c.bchain] : 0.0;\n",
         "            " + T(3) + "\n            const double xc = (c.isB && c.bchain >= 0) ? L.C[c.bchain] : 0.0;\n"),
        ("            bool ok = sbz_wb_lu<SbzBcWave>(c.lane, row, rpo);\n", "            " + T(4) + "\n            bool ok = sbz_wb_lu<SbzBcWave>(c.lane, row, rpo);\n"),
        ("            sbz_wave_apply(c, B0, L.dB[0], L.Cs, k, Un, DT);  // ∂f/∂τ at (t, u_n)\n",
         "            " + T(5) + "\n            sbz_wave_apply(c, B0, L.dB[0], L.Cs, k, Un, DT);  // ∂f/∂τ at (t, u_n)\n"),
        ("            st_j++;\n            __syncthreads();\n", "            st_j++;\n            __syncthreads();\n            " + T(6) + "\n"),
        ("                        __syncthreads();\n                        sbz_wave_apply(c, L.B[sbz_stage_bg[s]], nullptr, L.Cs, k, in, Ks);\n",
         "                        __syncthreads();\n                        " + T(7) + "\n                        sbz_wave_apply(c, L.B[sbz_stage_bg[s]], nullptr, L.Cs, k, in, Ks);\n"),
        ("                        st_f++;\n                        __syncthreads();\n", "                        st_f++;\n                        __syncthreads();\n                        " + T(8) + "\n"),
        ("                    __syncthreads();\n                    // solve W x = Ks:", "                    __syncthreads();\n                    " + T(9) + "\n                    // solve W x = Ks:"),
        ("                        L.C[c.lane] = y;\n                    }\n                    __syncthreads();\n",
         "                        L.C[c.lane] = y;\n                    }\n                    __syncthreads();\n                    " + T(10) + "\n"),
        ("                            L.X[c.lane] = x;\n                        }\n                    }\n                    __syncthreads();\n",
         "                            L.X[c.lane] = x;\n                        }\n                    }\n                    __syncthreads();\n                    " + T(11) + "\n"),
        ("                        SBZ_CHAIN_GROUPS(SBZ_CG_BWD)\n#undef SBZ_CG_BWD\n                    }\n                    __syncthreads();\n",
         "                        SBZ_CHAIN_GROUPS(SBZ_CG_BWD)\n#undef SBZ_CG_BWD\n                    }\n                    __syncthreads();\n                    " + T(12) + "\n"),
        ("            __syncthreads();\n            // the controller's two powers", "            __syncthreads();\n            " + T(13) + "\n            // the controller's two powers"),
        ("            dt = dt / q;\n        }\n", "            dt = dt / q;\n            " + T(14) + "\n        }\n"),
        ("            st[5] = 8 * st_j;\n        }\n", "            st[5] = 8 * st_j;\n        }\n        if (c.lane == 0) { double* o = a.out_u + (size_t)inst * SBZ_N;"
         f" for (int j = 0; j < {NR}; ++j) o[j] = (double)PR[j]; o[SBZ_N - 1] = -(double)iters; }}\n"),
    ])
    return src[:i0] + body + src[i1:]


def instrumented_model(M, kernel):
    """A GpuModel of M running the instrumented copy of `kernel` ("wave4" or "wave")."""
    from .solver import GpuModel
    em = model_source(M)
    src = wave4_source(em.full_src) if kernel == "wave4" else wave_source(em.full_src)
    path = KCACHE / f"regions_{kernel}.hsaco"
    compile_code_object(src, path)
    saved = em.full_hash
    em.full_hash = path.stem   # GpuModel loads _kcache/<full_hash>.hsaco
    try:
        return GpuModel(M)
    finally:
        em.full_hash = saved


def main(argv=None):
    from .constants import k_dimensionless
    from .problem import CosmologyProblem, solvebg, solvebg_sets
    from .solver import ALGO_WAVE, ALGO_WAVE4, SAVE_SOURCES
    from . import workloads
    ap = argparse.ArgumentParser(description=__doc__.splitlines()[0])
    ap.add_argument("kernel", choices=["wave4", "wave"])
    ap.add_argument("--sets", type=int, default=32, help="wave4: Latin-hypercube sets of the 128-set batch")
    a = ap.parse_args(argv)
    if a.kernel == "wave4":
        M, sets = workloads.lhs_sets(128, take=a.sets)
        ks_mpc = np.logspace(-5, 1, 512)
        bgs = solvebg_sets([CosmologyProblem(M, p) for p in sets])
        gm = instrumented_model(M, "wave4")
        iset, ks, t0 = [], [], []
        for si, (p, bg) in enumerate(zip(sets, bgs)):
            gm.upload_background(si, bg.t, bg.u, bg.du, bg.P)
            k = k_dimensionless(ks_mpc, p["h"])
            iset.append(np.full(len(k), si, np.int32))
            ks.append(k)
            t0.append(np.clip(np.minimum(1e-2 / k, 1e-4), bg.t[0], bg.t[-1]))
        plan = gm.plan(np.concatenate(iset), np.concatenate(ks), np.concatenate(t0), algo=ALGO_WAVE4)
        names, what = WAVE4_REGIONS, f"sbz_wave4_kernel, {a.sets} Latin-hypercube sets x 512 k (n = 38)"
    else:
        M, sets, ks_mpc, _ = workloads.workload("massive_nu100")
        p = sets[0]
        bg = solvebg(CosmologyProblem(M, p))
        gm = instrumented_model(M, "wave")
        gm.upload_background(0, bg.t, bg.u, bg.du, bg.P)
        k = k_dimensionless(ks_mpc, p["h"])
        t0 = np.clip(np.minimum(1e-2 / k, 1e-4), bg.t[0], bg.t[-1])
        taus = workloads.source_taus(bg.t[0], bg.t[-1], 2000)
        plan = gm.plan(np.zeros(len(k), np.int32), k, t0, saveat=taus, save=SAVE_SOURCES, algo=ALGO_WAVE)
        names, what = WAVE_REGIONS, "sbz_wave_kernel, n = 100, 1024 k, sources on 2000 tau"
    plan.run()
    ms = plan.run()
    U = plan.fetch()["u"]
    rows = U[:, -1] < 0
    if a.kernel == "wave":   # the latency-bound launch: the longest mode sets the kernel time
        i = int(np.argmax(-U[:, -1]))
        S, it, label = U[i, :len(names)], -U[i, -1], "the longest mode"
    else:
        S, it, label = U[rows, :len(names)].sum(axis=0), -U[rows, -1].sum(), "all wavefronts"
    print(f"{what}: kernel {ms:.1f} ms; regions over {label} ({it:.0f} step attempts)")
    tot = S.sum()
    for j, nm in enumerate(names):
        print(f"  {nm:22s} {S[j] / tot:6.3f}  {S[j] / it:9.0f} clock/attempt")
    print(f"  {'total':22s} {1.0:6.3f}  {tot / it:9.0f} clock/attempt")
    plan.close()
    gm.close()


if __name__ == "__main__":
    main()
