"""Kernel-variant autotuner for the batched perturbation solve.

For a plan it picks among the wavefront kernels:
- SBZ_ALGO_WAVE: one mode per wavefront, one resident wavefront per SIMD, every register
  available: the lowest latency per mode;
- SBZ_ALGO_WAVE_ALTLU: the same with the other LDS layout of the border factors (split rows with
  unmasked solves, or combined rows with half the LDS; the model's default is split unless that
  costs resident workgroups): fewer instructions per step against more modes in flight;
- SBZ_ALGO_WAVE2: one mode per wavefront, two resident per SIMD: the second hides the first's
  latency once there are more instances than SIMDs;
- SBZ_ALGO_WAVE4: four modes per wavefront (16 lanes each), for models whose border and chains
  fit 16 lanes and plans that save nothing: the per-step work that the lanes of a mode share
  (border RHS and Jacobian assembly, background) serves four modes;
- SBZ_ALGO_WAVE2H: two modes per wavefront (32 lanes each), the same for the models whose border
  or chains exceed 16 lanes but fit 32.

When the instances fit the one-per-SIMD capacity the launch is the latency of its longest mode:
the two factor layouts are timed on the 64 longest instances (when the other layout holds them
all at once) and the faster kept.  Otherwise every variant is timed on a cost-stratified sample
of the plan's own instances (every m-th of the instances sorted by k) large enough (8 × the
largest capacity, or all) that its time is set by throughput rather than by its single longest
mode, and the fastest is kept.  Choices are cached per model, instance-count bucket and output
kind.
"""
import ctypes
import math

import numpy as np

from . import _lib

ALGO_WAVE, ALGO_WAVE2, ALGO_WAVE4, ALGO_WAVE_ALTLU, ALGO_WAVE2H = 1, 3, 4, 5, 6
NAMES = {ALGO_WAVE: "WAVE", ALGO_WAVE2: "WAVE2", ALGO_WAVE4: "WAVE4", ALGO_WAVE_ALTLU: "WAVE_ALTLU",
         ALGO_WAVE2H: "WAVE2H"}
_cache = {}
last_times = {}


def capacity(gm, algo):
    """Resident modes of a variant on the model's device (0 when the model has no such kernel)."""
    L = gm.L
    L.sbz_model_capacity.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.POINTER(ctypes.c_int)]
    c = ctypes.c_int(0)
    _lib.check(L.sbz_model_capacity(gm.h, algo, ctypes.byref(c)), "sbz_model_capacity")
    return c.value


def choose(gm, inst_set, ks, t0, rtol=1e-5, atol=1e-5, maxiters=100000, saveat=None, save=0, verbose=False):
    """The kernel variant for this batch (see module docstring); returns an SBZ_ALGO_* value."""
    n = len(ks)
    n_save = 0 if saveat is None else len(saveat)
    key = (gm.em.full_hash, int(math.log2(max(n, 1))), save, n_save > 0)
    if key in _cache:
        return _cache[key]
    caps = {ALGO_WAVE: capacity(gm, ALGO_WAVE), ALGO_WAVE2: capacity(gm, ALGO_WAVE2)}
    c_alt = capacity(gm, ALGO_WAVE_ALTLU)
    if c_alt > 0:
        caps[ALGO_WAVE_ALTLU] = c_alt
    if n_save == 0:
        for al in (ALGO_WAVE4, ALGO_WAVE2H):
            c = capacity(gm, al)
            if c > 0:
                caps[al] = c
    order = np.argsort(-np.asarray(ks), kind="stable")
    if n <= caps[ALGO_WAVE] or max(caps.values()) <= caps[ALGO_WAVE]:
        # latency-bound: the longest modes set the launch; only the factor layout can matter
        cands = [ALGO_WAVE] + ([ALGO_WAVE_ALTLU] if caps.get(ALGO_WAVE_ALTLU, 0) >= n else [])
        if len(cands) == 1:
            _cache[key] = ALGO_WAVE
            return ALGO_WAVE
        sel = np.sort(order[:min(n, 64)])
    else:
        cands = list(caps)
        m = max(1, n // (8 * max(caps.values())))
        sel = np.sort(order[::m])
    times = {}
    for algo in cands:
        plan = gm.plan(np.asarray(inst_set)[sel], np.asarray(ks)[sel], np.asarray(t0)[sel], rtol=rtol, atol=atol,
                       maxiters=maxiters, saveat=saveat, save=save, algo=algo)
        times[algo] = plan.run()
        plan.close()
    best = min(times, key=times.get)
    last_times.clear()
    last_times.update({NAMES[a]: t for a, t in times.items()})
    if verbose:
        print(f"autotune: {len(sel)} of {n} instances; " + ", ".join(f"{NAMES[a]} {t:.1f} ms" for a, t in times.items())
              + f" -> {NAMES[best]}")
    _cache[key] = best
    return best
