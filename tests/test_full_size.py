"""The benchmark workloads at their full sizes on the GPU, through properties that do not need an
oracle run of every instance: every mode succeeds; the results of an instance do not depend on
where it sits in the batch (the batch reversed gives bit-identical states, step counts and P(k));
and a random sample of instances matches the oracle at 1e-4 (P(k), the parity bar of the smaller
tests).  Workloads: the 128-set Latin-hypercube batch (65,536 instances, n = 38), the n = 100
massive-ν model over 64 of those sets × 1024 k (65,536 instances), and the n = 100 model at 1024 k
with CMB sources on 2000 τ."""
import sys

import numpy as np
import pytest

from conftest import Setup
from symboltz import workloads
from symboltz.constants import k_dimensionless
from symboltz.problem import CosmologyProblem, solvebg, solvebg_sets
from symboltz.solver import GpuModel, ALGO_AUTO, ALGO_WAVE, SAVE_SOURCES

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _instances(M, sets, bgs, ks_mpc, gm):
    iset, ks, t0 = [], [], []
    for si, (p, bg) in enumerate(zip(sets, bgs)):
        gm.upload_background(si, bg.t, bg.u, bg.du, bg.P)
        k = k_dimensionless(ks_mpc, p["h"])
        iset.append(np.full(len(k), si, np.int32))
        ks.append(k)
        t0.append(np.clip(np.minimum(1e-2 / k, 1e-4), bg.t[0], bg.t[-1]))
    return np.concatenate(iset), np.concatenate(ks), np.concatenate(t0)


@pytest.mark.timeout(600)
def test_latin_hypercube_batch_full_size():
    M, sets, ks_mpc, _ = workloads.workload("lhs128")
    say(f"case: {len(sets)} sets x {len(ks_mpc)} k = {len(sets) * len(ks_mpc)} instances, n = {M.n_unknowns}")
    bgs = solvebg_sets([CosmologyProblem(M, p) for p in sets])
    gm = GpuModel(M)
    iset, ks, t0 = _instances(M, sets, bgs, ks_mpc, gm)
    plan = gm.plan(iset, ks, t0, algo=ALGO_AUTO)
    say(f"kernel variant {plan.algo}: {plan.run():.1f} ms")
    r = plan.fetch()
    pk = plan.pk()
    assert np.all(r["retcode"] == 0)
    assert np.all(np.isfinite(pk)) and np.all(pk > 0)
    # the batch reversed: each instance's results are bit for bit the same
    rev = np.arange(len(ks))[::-1].copy()
    plan2 = gm.plan(iset[rev], ks[rev], t0[rev], algo=plan.algo)
    plan2.run()
    r2 = plan2.fetch()
    np.testing.assert_array_equal(r2["u"][::-1], r["u"])
    np.testing.assert_array_equal(r2["stats"][::-1], r["stats"])
    np.testing.assert_array_equal(plan2.pk()[::-1], pk)
    # a random sample against the oracle
    rng = np.random.default_rng(7)
    pick = rng.choice(len(ks), 48, replace=False)
    worst = 0.0
    for si in np.unique(iset[pick]):
        m = pick[iset[pick] == si]
        S = Setup(M, sets[si])
        P, ro = S.O.spectrum_matter(ks[m])
        assert np.all(ro["retcode"] == 0)
        worst = max(worst, float(np.max(np.abs(pk[m] / P - 1))))
    say(f"sample of {len(pick)} instances: max |P/P_oracle - 1| = {worst:.2e}")
    assert worst < 1e-4
    plan.close()
    plan2.close()
    gm.close()


@pytest.mark.timeout(600)
def test_massive_nu_batched_full_size():
    # the n = 100 model over 64 sets x 1024 k: the autotuner's variant (two modes per wavefront)
    M, sets, ks_mpc, _ = workloads.workload("massive_nu100_lhs64")
    say(f"case: {len(sets)} sets x {len(ks_mpc)} k = {len(sets) * len(ks_mpc)} instances, n = {M.n_unknowns}")
    bgs = solvebg_sets([CosmologyProblem(M, p) for p in sets])
    gm = GpuModel(M)
    iset, ks, t0 = _instances(M, sets, bgs, ks_mpc, gm)
    plan = gm.plan(iset, ks, t0, algo=ALGO_AUTO)
    say(f"kernel variant {plan.algo}: {plan.run():.1f} ms")
    r = plan.fetch()
    pk = plan.pk()
    assert np.all(r["retcode"] == 0)
    assert np.all(np.isfinite(pk)) and np.all(pk > 0)
    rev = np.arange(len(ks))[::-1].copy()
    plan2 = gm.plan(iset[rev], ks[rev], t0[rev], algo=plan.algo)
    plan2.run()
    r2 = plan2.fetch()
    np.testing.assert_array_equal(r2["u"][::-1], r["u"])
    np.testing.assert_array_equal(r2["stats"][::-1], r["stats"])
    np.testing.assert_array_equal(plan2.pk()[::-1], pk)
    rng = np.random.default_rng(11)
    pick = rng.choice(len(ks), 12, replace=False)
    worst = 0.0
    for si in np.unique(iset[pick]):
        m = pick[iset[pick] == si]
        S = Setup(M, sets[si])
        P, ro = S.O.spectrum_matter(ks[m])
        assert np.all(ro["retcode"] == 0)
        worst = max(worst, float(np.max(np.abs(pk[m] / P - 1))))
    say(f"sample of {len(pick)} instances: max |P/P_oracle - 1| = {worst:.2e}")
    assert worst < 1e-4
    plan.close()
    plan2.close()
    gm.close()


@pytest.mark.timeout(600)
def test_massive_nu_sources_full_size():
    M, sets, ks_mpc, _ = workloads.workload("massive_nu100")
    p = sets[0]
    bg = solvebg(CosmologyProblem(M, p))
    taus = workloads.source_taus(bg.t[0], bg.t[-1], 2000)
    say(f"case: n = {M.n_unknowns}, {len(ks_mpc)} k, sources on {len(taus)} tau")
    gm = GpuModel(M)
    iset, ks, t0 = _instances(M, [p], [bg], ks_mpc, gm)
    plan = gm.plan(iset, ks, t0, saveat=taus, save=SAVE_SOURCES, algo=ALGO_WAVE)
    say(f"{plan.run():.1f} ms")
    r = plan.fetch()
    src = plan.sources()
    pk = plan.pk()
    assert np.all(r["retcode"] == 0)
    assert np.all(np.isfinite(src)) and np.all(np.isfinite(pk)) and np.all(pk > 0)
    rev = np.arange(len(ks))[::-1].copy()
    plan2 = gm.plan(iset[rev], ks[rev], t0[rev], saveat=taus, save=SAVE_SOURCES, algo=ALGO_WAVE)
    plan2.run()
    np.testing.assert_array_equal(plan2.fetch()["u"][::-1], r["u"])
    np.testing.assert_array_equal(plan2.sources()[::-1], src)
    # sample against the oracle: P(k) at 1e-4; the sources where the late-time ISW term is
    # well conditioned (k <= 3000 H0/c; above it see test_configs.py)
    rng = np.random.default_rng(8)
    pick = np.sort(rng.choice(len(ks), 12, replace=False))
    S = Setup(M, p)
    o = S.O.solve(ks[pick], t0[pick], saveat=taus, sources=True)
    assert np.all(o["retcode"] == 0)
    P = np.array([S.O.p0(k) * S.O.delta_m(k, bg.t[-1], o["u"][i]) ** 2 for i, k in enumerate(ks[pick])])
    errp = float(np.max(np.abs(pk[pick] / P - 1)))
    lo = ks[pick] <= 3000.0
    sc = np.abs(o["src"][lo]).max(axis=1, keepdims=True) + 1e-300
    errs = float(np.max(np.abs(src[pick][lo] - o["src"][lo]) / sc)) if lo.any() else 0.0
    say(f"sample of {len(pick)} modes: P(k) {errp:.2e}, sources (k <= 3000 H0/c) {errs:.2e}")
    assert errp < 1e-4 and errs < 1e-4
    plan.close()
    plan2.close()
    gm.close()
