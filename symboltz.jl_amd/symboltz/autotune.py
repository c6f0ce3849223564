"""Kernel-variant autotuner for the batched perturbation solve.

For a plan it picks the wavefront kernel's occupancy variant: SBZ_ALGO_WAVE (one resident
wavefront per SIMD, every register available: the lowest latency per mode) or SBZ_ALGO_WAVE2 (two
per SIMD: the second hides the first's latency once there are more instances than SIMDs).  When
the instances fit the one-per-SIMD capacity the choice is WAVE without measuring; otherwise each
candidate is timed on a cost-stratified sample of the plan's own instances (every m-th of the
instances sorted by k) large enough (8 × the larger capacity, or all) that its time is set by
throughput rather than by its single longest mode, and the faster is kept.  Choices are
cached per model, instance-count bucket and output kind.
"""
import ctypes
import math

import numpy as np

from . import _lib

ALGO_WAVE, ALGO_LANE, ALGO_WAVE2 = 1, 2, 3
_cache = {}


def capacity(gm, algo):
    """Resident wavefronts (lanes for the lane kernel) of a variant on the model's device."""
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
    cap1, cap2 = capacity(gm, ALGO_WAVE), capacity(gm, ALGO_WAVE2)
    if n <= cap1 or cap2 <= cap1:
        _cache[key] = ALGO_WAVE
        return ALGO_WAVE
    order = np.argsort(-np.asarray(ks), kind="stable")
    m = max(1, n // (8 * cap2))
    sel = np.sort(order[::m])
    times = {}
    for algo in (ALGO_WAVE, ALGO_WAVE2):
        plan = gm.plan(np.asarray(inst_set)[sel], np.asarray(ks)[sel], np.asarray(t0)[sel], rtol=rtol, atol=atol,
                       maxiters=maxiters, saveat=saveat, save=save, algo=algo)
        times[algo] = plan.run()
        plan.close()
    best = min(times, key=times.get)
    if verbose:
        print(f"autotune: {len(sel)} of {n} instances; WAVE {times[ALGO_WAVE]:.1f} ms, WAVE2 "
              f"{times[ALGO_WAVE2]:.1f} ms -> {'WAVE2' if best == ALGO_WAVE2 else 'WAVE'}")
    _cache[key] = best
    return best
