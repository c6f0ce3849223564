"""The mixed-model request stream: spectra for requests across several model families.

Each request names a model family and its parameters.  The learned surrogate answers the requests
of the family it was trained on when they lie inside its training box and its members agree
within `tol` in log P; every other request (other families, out of the box, or disputed) is solved
by the plain GPU solver, and all those solves, of every family and state size, run in one
persistent launch (`symboltz.multi`): one queue, longest modes first across the families, one
tail for the whole batch.

No reference counterpart: the reference solves one cosmology of one model per call; this is the
batched form of calling `solve(prob, ks)` and `spectrum_matter` for every request.
"""
import time

import numpy as np

from .constants import k_dimensionless
from .multi import MultiModel
from .problem import CosmologyProblem, solvebg_sets
from .solver import GpuModel


class MixedStream:
    """families: model descriptions (at most 8); ks_mpc: the k grid of every spectrum (1/Mpc).
    surrogate / surrogate_family: a trained `symboltz.surrogate.Surrogate` on this k grid and the
    index of the family whose requests it may answer (None: every request is solved)."""

    def __init__(self, families, ks_mpc, surrogate=None, surrogate_family=None, tol=2e-3, rtol=1e-5, atol=1e-5):
        self.families = list(families)
        self.ks_mpc = np.asarray(ks_mpc, dtype=np.float64)
        if surrogate is not None:
            if surrogate.widths[-1] != len(self.ks_mpc):
                raise ValueError(f"the surrogate predicts {surrogate.widths[-1]} k, the stream's grid has {len(self.ks_mpc)}")
            if surrogate_family is None or not 0 <= surrogate_family < len(self.families):
                raise ValueError("a surrogate needs the index of the family it was trained on")
        self.surrogate = surrogate
        self.sfam = surrogate_family
        self.tol = tol
        self.rtol, self.atol = rtol, atol
        self.gms = [GpuModel(M) for M in self.families]
        self.multi = MultiModel(self.gms)

    def _plans(self, groups):
        """One plan per family over its requests' sets × k (sets uploaded as 0..m-1)."""
        plans, shapes = [], []
        for f, pars in enumerate(groups):
            gm, M = self.gms[f], self.families[f]
            bgs = solvebg_sets([CosmologyProblem(M, p) for p in pars]) if pars else []
            iset, kk, tt = [], [], []
            for s, (p, bg) in enumerate(zip(pars, bgs)):
                gm.upload_background(s, bg.t, bg.u, bg.du, bg.P)
                k = k_dimensionless(self.ks_mpc, p["h"])
                iset.append(np.full(len(k), s, np.int32))
                kk.append(k)
                tt.append(np.clip(np.minimum(1e-2 / k, 1e-4), bg.t[0], bg.t[-1]))
            cat = (lambda xs, dt: np.concatenate(xs) if xs else np.zeros(0, dt))
            plans.append(gm.plan(cat(iset, np.int32), cat(kk, np.float64), cat(tt, np.float64),
                                 rtol=self.rtol, atol=self.atol))
            shapes.append(len(pars))
        return plans, shapes

    def solve(self, requests, mode="one_launch"):
        """requests: (family index, parameter dict, surrogate input row or None) per request.
        Returns log P [n_req][n_k] (P in (c/H0)³), the mask of requests the surrogate answered,
        the per-request return code (0: every mode succeeded) and timings in ms (surrogate
        inference, background solves and uploads, solve kernel(s), P(k)), with the step attempts
        and their algorithmic flops of the solves.  `mode` picks how the solves run:
        "one_launch" (sbz_multi_kernel over all families), "concurrent" (one launch per family, each
        on its model's stream, all enqueued before any is waited for: the kernels overlap on the
        device, each at its own occupancy) or "serial" (one launch per family, each waited for).
        solve_ms is the one launch's kernel time, the longest of the concurrent launches from the
        first enqueue (wall clock), or the sum of the serial ones."""
        n, nk = len(requests), len(self.ks_mpc)
        logp = np.full((n, nk), np.nan)
        answered = np.zeros(n, bool)
        ret = np.zeros(n, np.int32)
        tm = {"surrogate_ms": 0.0, "prepare_ms": 0.0, "solve_ms": 0.0, "pk_ms": 0.0}
        if self.surrogate is not None and self.sfam is not None:
            idx = [i for i, (f, _, x) in enumerate(requests) if f == self.sfam and x is not None]
            if idx:
                X = np.array([requests[i][2] for i in idx], dtype=np.float64)
                mean, dis, inbox, ms = self.surrogate.predict(X)
                tm["surrogate_ms"] = ms
                ok = inbox & (dis <= self.tol)
                for j, i in enumerate(idx):
                    if ok[j]:
                        logp[i] = mean[j]
                        answered[i] = True
        groups = [[] for _ in self.families]
        where = [[] for _ in self.families]
        for i, (f, p, _) in enumerate(requests):
            if not answered[i]:
                groups[f].append(p)
                where[f].append(i)
        t0 = time.perf_counter()
        plans, _ = self._plans(groups)
        tm["prepare_ms"] = (time.perf_counter() - t0) * 1e3
        try:
            if mode == "one_launch":
                tm["solve_ms"] = self.multi.run(plans)
            elif mode == "concurrent":
                live = [p for p in plans if p.n_inst]
                t1 = time.perf_counter()
                for p in live:
                    p.launch()
                for p in live:
                    p.wait()
                tm["solve_ms"] = (time.perf_counter() - t1) * 1e3 if live else 0.0
            elif mode == "serial":
                tm["solve_ms"] = sum(p.run() for p in plans if p.n_inst)
            else:
                raise ValueError(f"unknown mode {mode!r}")
            t0 = time.perf_counter()
            tm["steps"] = 0
            tm["flops"] = 0
            for f, p in enumerate(plans):
                if not p.n_inst:
                    continue
                pk = p.pk().reshape(len(where[f]), nk)
                r = p.fetch()
                rc = r["retcode"].reshape(len(where[f]), nk)
                steps = int(r["stats"][:, 0].sum() + r["stats"][:, 1].sum())
                tm["steps"] += steps
                tm["flops"] += steps * self.gms[f].em.flops_step   # the emitter's count per Rodas5P step
                for j, i in enumerate(where[f]):
                    logp[i] = np.log(pk[j])
                    ret[i] = int(np.max(rc[j]))
            tm["pk_ms"] = (time.perf_counter() - t0) * 1e3
        finally:
            for p in plans:
                p.close()
        return logp, answered, ret, tm

    def close(self):
        self.multi.close()
        for gm in self.gms:
            gm.close()
