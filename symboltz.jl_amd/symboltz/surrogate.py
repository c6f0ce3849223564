"""Learned surrogate for the matter power spectrum, with the plain GPU solver as fallback and judge.

A small ensemble of MLPs maps seven cosmological parameters to log P(k) on a fixed k grid.
- **Training data:** spectra the GPU solver itself computes.
- **Networks:** `symboltz.mlp`, with BF16 MFMA GEMMs, FP32 accumulation, FP32 master weights
  and Adam.
- **Fallback:** a request is answered by the plain solver instead when its parameters lie outside
  the training box, or when the members disagree by more than a tolerance in log P.
- **Judge:** the plain solver measures the surrogate's error on held-out parameter sets.

No reference counterpart: this is the learned part of the build's design.

Parameters (the batched configuration's ranges plus the summed neutrino mass):
- h, ω_b = Ω_b h², ω_c = Ω_c h², ln(10¹⁰ A_s), n_s;
- the hydrogen reionization redshift z_re (the solver's parameter; a τ_reion input would add a
  root-find per set, see `problem.z_re_for_optical_depth`);
- Σm_ν, as three degenerate massive species of Σ/3 each.
"""
import math
import time

import numpy as np

from . import models
from .constants import k_dimensionless
from .mlp import Mlp
from .problem import CosmologyProblem, solvebg_sets

PARAMS = ["h", "omega_b", "omega_c", "ln_As1e10", "ns", "z_re", "Sum_mnu"]
LO = np.array([0.60, 0.020, 0.10, 2.9, 0.92, 6.2, 0.06])
HI = np.array([0.75, 0.024, 0.14, 3.2, 1.00, 10.3, 0.30])


def surrogate_model():
    """ΛCDM with three degenerate massive neutrinos (4 momentum bins), hierarchies truncated at 6."""
    return models.LCDM(lmax=6)


def k_grid(n=256):
    """n k-values log-spaced in [1e-4, 5]/Mpc."""
    return np.logspace(-4, math.log10(5.0), n)


def lhs(n, seed):
    """n Latin-hypercube points in the parameter box (rows: PARAMS order)."""
    from scipy.stats import qmc
    return qmc.scale(qmc.LatinHypercube(d=len(PARAMS), seed=seed).random(n), LO, HI)


def pars_of(M, x):
    h, wb, wc, lnAs, ns, zre, smnu = [float(v) for v in x]
    p = dict(models.parameters_Planck18(M))
    p.update(h=h, Omega_b0=wb / h**2, Omega_c0=wc / h**2, ln_As1e10=lnAs, ns=ns, z_re=zre, N=3, m_eV=smnu / 3)
    return p


def solve_spectra(X, ks_mpc=None, M=None, chunk=1024, algo=None, log=None):
    """log P(k) [n][n_k] for every parameter row of X with the plain GPU solver (P in (c/H0)³ units),
    chunk sets per launch; returns (logP, seconds of GPU kernel time, seconds wall)."""
    from .solver import ALGO_AUTO, GpuModel
    M = M or surrogate_model()
    ks_mpc = k_grid() if ks_mpc is None else np.asarray(ks_mpc)
    gm = GpuModel(M)
    out = np.zeros((len(X), len(ks_mpc)))
    kern = 0.0
    t_wall = time.perf_counter()
    for c0 in range(0, len(X), chunk):
        rows = X[c0:c0 + chunk]
        pars = [pars_of(M, x) for x in rows]
        # backgrounds of the chunk: one GPU launch from 256 sets on, host threads below
        bgs = solvebg_sets([CosmologyProblem(M, p) for p in pars])
        iset, kk, tt = [], [], []
        for s, (p, bg) in enumerate(zip(pars, bgs)):
            gm.upload_background(s, bg.t, bg.u, bg.du, bg.P)
            k = k_dimensionless(ks_mpc, p["h"])
            iset.append(np.full(len(k), s, np.int32))
            kk.append(k)
            tt.append(np.clip(np.minimum(1e-2 / k, 1e-4), bg.t[0], bg.t[-1]))
        plan = gm.plan(np.concatenate(iset), np.concatenate(kk), np.concatenate(tt),
                       algo=ALGO_AUTO if algo is None else algo)
        if log and c0 == 0:
            from . import autotune
            log(f"kernel variant {autotune.NAMES.get(plan.algo, plan.algo)}"
                + (f" (sample ms: {autotune.last_times})" if algo is None and autotune.last_times else ""))
        kern += plan.run() * 1e-3
        r = plan.fetch()
        if not np.all(r["retcode"] == 0):
            raise RuntimeError("a perturbation solve failed while generating spectra")
        out[c0:c0 + len(rows)] = np.log(plan.pk().reshape(len(rows), len(ks_mpc)))
        plan.close()
        if log:
            log(f"spectra {c0 + len(rows)}/{len(X)}")
    return out, kern, time.perf_counter() - t_wall


class Surrogate:
    """Ensemble of MLPs on standardized inputs (the box mapped to [-1, 1]) and outputs (log P minus
    its training mean, over its training standard deviation, per k)."""

    def __init__(self, n_members=2, hidden=(512, 512, 512, 512), n_out=256, max_batch=65536, seed=0):
        self.widths = [len(PARAMS), *hidden, n_out]
        self.nets = [Mlp(self.widths, max_batch, seed=seed + 1000 * i) for i in range(n_members)]
        self.mu = self.sd = None

    def _xin(self, X):
        return (2.0 * (np.asarray(X) - LO) / (HI - LO) - 1.0).astype(np.float32)

    def fit(self, X, logP, epochs=200, batch=1024, lr=1e-3, seed=0, log=None):
        """Adam over `epochs` passes (learning rate cosine-decayed to 2 % of lr); returns the
        per-epoch mean training loss of every member and the GPU training time."""
        self.mu = logP.mean(axis=0)
        self.sd = logP.std(axis=0) + 1e-12
        T = ((logP - self.mu) / self.sd).astype(np.float32)
        Xi = self._xin(X)
        hist, ms = [], 0.0
        for i, net in enumerate(self.nets):
            net.set_data(Xi, T)
            rng = np.random.default_rng(seed + i)
            h = []
            for ep in range(epochs):
                lr_e = lr * (0.02 + 0.98 * 0.5 * (1 + math.cos(math.pi * ep / epochs)))
                perm = rng.permutation(len(Xi))[: (len(Xi) // batch) * batch]
                loss, t = net.train(perm, batch, lr_e)
                ms += t
                h.append(float(loss.mean()))
                if log and (ep % 50 == 0 or ep == epochs - 1):
                    log(f"member {i} epoch {ep}: loss {h[-1]:.3e}")
            hist.append(h)
        return hist, ms * 1e-3

    def predict(self, X):
        """(log P mean over members, max |member - mean| in log P per row, in-box mask, kernel ms)."""
        Xi = self._xin(X)
        outs, ms = [], 0.0
        for net in self.nets:
            y, t = net.forward(Xi)
            outs.append(y.astype(np.float64) * self.sd + self.mu)
            ms += t
        outs = np.array(outs)
        mean = outs.mean(axis=0)
        disagree = np.abs(outs - mean).max(axis=(0, 2))
        inbox = np.all((np.asarray(X) >= LO) & (np.asarray(X) <= HI), axis=1)
        return mean, disagree, inbox, ms

    def predict_with_fallback(self, X, tol=2e-3, ks_mpc=None):
        """log P for every row: the surrogate where it is trusted (inside the box and members
        agreeing within tol in log P), the plain GPU solver elsewhere; returns (log P, fallback mask)."""
        mean, disagree, inbox, _ = self.predict(X)
        fb = ~inbox | (disagree > tol)
        if np.any(fb):
            mean[fb] = solve_spectra(np.asarray(X)[fb], ks_mpc)[0]
        return mean, fb

    def close(self):
        for n in self.nets:
            n.close()


def main():
    """End-to-end run: generate training and held-out spectra with the GPU solver, train the
    ensemble, judge it on the held-out sets, and time inference against the plain solver."""
    import argparse
    import json
    import sys
    ap = argparse.ArgumentParser()
    ap.add_argument("--train", type=int, default=8192)
    ap.add_argument("--heldout", type=int, default=1024)
    ap.add_argument("--epochs", type=int, default=300)
    ap.add_argument("--members", type=int, default=2)
    ap.add_argument("--data", default=None, help="npz with X, logP, Xh, logPh (skips the generation)")
    ap.add_argument("--save-data", default=None)
    ap.add_argument("--infer-batch", type=int, default=65536)
    ap.add_argument("--batch", type=int, default=1024)
    ap.add_argument("--lr", type=float, default=1e-3)
    ap.add_argument("--generate-only", action="store_true")
    ap.add_argument("--tol", type=float, default=5e-3, help="fallback threshold on member disagreement in log P")
    ap.add_argument("--stream", type=int, default=0,
                    help="then run a mixed-model request stream of this many requests with the trained surrogate")
    ap.add_argument("--stream-out-of-box", type=float, default=0.1,
                    help="fraction of the surrogate family's stream requests placed outside its box")
    args = ap.parse_args()
    log = lambda *a: print(*a, file=sys.stderr, flush=True)   # noqa: E731
    if args.data:
        d = np.load(args.data)
        X, Y, Xh, Yh = (d[k].astype(np.float64) for k in ("X", "logP", "Xh", "logPh"))
        gen_s = {k: float(d[k]) for k in ("train_kernel_s", "train_wall_s", "heldout_kernel_s", "heldout_wall_s")
                 if k in d.files} or None
    else:
        X, Xh = lhs(args.train, seed=1), lhs(args.heldout, seed=2)
        Y, gk, gw = solve_spectra(X, log=log)
        Yh, hk, hw = solve_spectra(Xh, log=log)
        gen_s = dict(train_kernel_s=gk, train_wall_s=gw, heldout_kernel_s=hk, heldout_wall_s=hw)
        if args.save_data:
            np.savez(args.save_data, X=X.astype(np.float32), logP=Y.astype(np.float32), Xh=Xh.astype(np.float32),
                     logPh=Yh.astype(np.float32), **{k: np.array(v) for k, v in gen_s.items()})
        if args.generate_only:
            print(json.dumps(dict(gen_s, train_sets=len(X), heldout_sets=len(Xh))), flush=True)
            return
    S = Surrogate(n_members=args.members)
    hist, train_s = S.fit(X, Y, epochs=args.epochs, batch=args.batch, lr=args.lr, log=log)
    mean, disagree, inbox, _ = S.predict(Xh)
    err = np.abs(np.exp(mean - Yh) - 1.0).max(axis=1)          # max |P_sur/P_solver - 1| per set
    # inference throughput at the requested batch (parameters drawn in the box)
    Xb = lhs(args.infer_batch, seed=3)
    S.predict(Xb[:1024])
    _, _, _, ms = S.predict(Xb)
    res = {
        "train_sets": len(X), "heldout_sets": len(Xh), "n_k": Y.shape[1], "members": args.members,
        "widths": S.widths, "epochs": args.epochs, "final_loss": [h[-1] for h in hist],
        "gpu_train_s": train_s, "generation": gen_s,
        "heldout_max_rel_err": float(err.max()), "heldout_p99_rel_err": float(np.quantile(err, 0.99)),
        "heldout_median_rel_err": float(np.median(err)), "heldout_disagreement_max": float(disagree.max()),
        "infer_batch": args.infer_batch, "infer_ms_all_members": ms,
        "surrogate_spectra_per_s": args.infer_batch / (ms * 1e-3),
    }
    # with fallback: requests outside the box or with disagreement above tol go to the plain
    # solver, the reference the errors are measured against (so their error is zero); out-of-box
    # requests are exercised on 8 shifted sets
    fb = ~inbox | (disagree > args.tol)
    err_fb = np.where(fb, 0.0, err)
    Xo = Xh[:8].copy()
    Xo[:, 6] = HI[6] * 1.5                                      # Σm_ν beyond the box
    yo, fbo = S.predict_with_fallback(Xo, tol=args.tol)
    yd = solve_spectra(Xo)[0]
    res.update(fallback_tol_logP=args.tol, heldout_fallback_fraction=float(fb.mean()),
               heldout_max_rel_err_with_fallback=float(err_fb.max()),
               out_of_box_fallback=bool(np.all(fbo)), out_of_box_max_rel_err=float(np.abs(np.exp(yo - yd) - 1).max()))
    if gen_s:
        res["plain_spectra_per_s"] = len(Xh) / gen_s["heldout_kernel_s"]
        res["speedup_vs_plain"] = res["surrogate_spectra_per_s"] / res["plain_spectra_per_s"]
        # a batch with the held-out fallback fraction: surrogate for all, plain solves for the fallbacks
        t_eff = args.infer_batch / res["surrogate_spectra_per_s"] + fb.mean() * args.infer_batch / res["plain_spectra_per_s"]
        res["with_fallback_spectra_per_s"] = args.infer_batch / t_eff
        res["with_fallback_speedup_vs_plain"] = res["with_fallback_spectra_per_s"] / res["plain_spectra_per_s"]
    if args.stream:
        res["stream"] = stream_run(S, args.stream, args.stream_out_of_box, args.tol, log)
    print(json.dumps(res), flush=True)
    S.close()


def stream_run(S, n, out_of_box, tol, log):
    """The mixed-model request stream with this surrogate: n requests over the five families of
    `workloads.mixed_families` (256 k each), the surrogate answering the in-box, agreed requests of
    its family and one persistent launch solving every other request; then the answered requests
    solved by the plain solver, the judge of the surrogate's error."""
    from . import workloads
    from .stream import MixedStream
    fams = workloads.mixed_families()
    sf = workloads.MIXED_SURROGATE_FAMILY
    if fams[sf] != surrogate_model():
        raise RuntimeError("the stream's surrogate family is not the surrogate's model")
    ks = k_grid(S.widths[-1])
    reqs = workloads.mixed_requests(n, seed=0, out_of_box=out_of_box)
    st = MixedStream(fams, ks, surrogate=S, surrogate_family=sf, tol=tol)
    plain = MixedStream(fams, ks)
    try:
        st.solve(reqs[:2 * len(fams)])                          # module loads and first-call costs
        t0 = time.perf_counter()
        lp, ans, ret, tm = st.solve(reqs)
        wall = time.perf_counter() - t0
        log(f"stream: {n} requests in {wall:.2f} s, {int(ans.sum())} answered by the surrogate; {tm}")
        idx = np.flatnonzero(ans)
        err = 0.0
        if len(idx):
            lp0, _, ret0, _ = plain.solve([reqs[i] for i in idx])
            if not np.all(ret0 == 0):
                raise RuntimeError("a plain solve of an answered request failed")
            err = float(np.abs(np.exp(lp[idx] - lp0) - 1.0).max())
        fam = np.array([r[0] for r in reqs])
        return {
            "requests": n, "families": [M.n_unknowns for M in fams], "n_k": len(ks),
            "surrogate_family": sf, "requests_of_surrogate_family": int((fam == sf).sum()),
            "answered_by_surrogate": int(ans.sum()),
            "out_of_box_requests": int(sum(1 for r in reqs if r[0] == sf and r[2] is not None
                                           and not np.all((r[2] >= LO) & (r[2] <= HI)))),
            "solved_modes": int((~ans).sum()) * len(ks), "all_success": bool(np.all(ret == 0)),
            "wall_s": wall, "spectra_per_s": n / wall, "stage_ms": {k: float(v) for k, v in tm.items()},
            "answered_max_rel_err_vs_solver": err, "tol_logP": tol,
        }
    finally:
        plain.close()
        st.close()


if __name__ == "__main__":
    main()
