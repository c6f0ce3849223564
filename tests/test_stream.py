"""The mixed-model stream on the GPU: five model families (38 to 236 unknowns) solved in one
persistent launch (sbz_multi_kernel), against the same plans run one launch per family and against
the oracle; and the routing of requests between the surrogate and the solver.  Bars: P(k) ≤ 1e-4
against the oracle (test/runtests.jl:872-876 tolerance form), the one-launch results identical to
the per-family launches (the same arithmetic per mode, whatever the schedule)."""
import sys

import numpy as np
import pytest

from conftest import Setup
from symboltz import workloads
from symboltz.constants import k_dimensionless

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def test_mixed_families_one_launch_match_per_family_and_oracle():
    from symboltz.stream import MixedStream
    fams = workloads.mixed_families()
    ks = np.logspace(-4, 0.5, 10)
    reqs = workloads.mixed_requests(10, seed=5)
    say(f"case: {len(reqs)} requests over {len(fams)} families (n = {[M.n_unknowns for M in fams]}), "
        f"{len(ks)} k, one launch vs one launch per family")
    st = MixedStream(fams, ks)
    try:
        lp1, ans1, ret1, tm1 = st.solve(reqs)
        lp2, ans2, ret2, tm2 = st.solve(reqs, mode="serial")
        lp5, _, ret5, tm5 = st.solve(reqs, mode="concurrent")
        say(f"one launch {tm1['solve_ms']:.1f} ms, per family serial {tm2['solve_ms']:.1f} ms, "
            f"concurrent {tm5['solve_ms']:.1f} ms")
        assert not ans1.any() and not ans2.any()
        assert np.all(ret1 == 0) and np.all(ret2 == 0) and np.all(ret5 == 0)
        np.testing.assert_array_equal(lp1, lp2)
        np.testing.assert_array_equal(lp1, lp5)
        # the oracle on the first request of every family
        for f in range(len(fams)):
            i = next(j for j, r in enumerate(reqs) if r[0] == f)
            M, p = fams[f], reqs[i][1]
            S = Setup(M, p)
            P, r = S.O.spectrum_matter(k_dimensionless(ks[::3], p["h"]))   # the oracle takes k in H0/c
            assert np.all(r["retcode"] == 0)
            err = np.max(np.abs(np.exp(lp1[i, ::3]) / P - 1))
            say(f"family {f} (n = {M.n_unknowns}): max |P/P_oracle - 1| = {err:.2e}")
            assert err < 1e-4, (f, err)
        # an empty family and a family-less batch
        lp3, _, ret3, _ = st.solve([r for r in reqs if r[0] != 2])
        np.testing.assert_array_equal(lp3, lp1[[j for j, r in enumerate(reqs) if r[0] != 2]])
        lp4, _, _, tm4 = st.solve([])
        assert lp4.shape == (0, len(ks)) and tm4["solve_ms"] == 0.0
    finally:
        st.close()


def test_mixed_launch_several_modes_bodies_match_per_family_and_oracle():
    """Family plans of at least twice the capacity run in their several-modes bodies inside the one
    launch (four modes per wavefront at n = 38, two at n = 54): against one launch per family with
    the one-mode kernel (P within 1e-6: the error norms' summation order differs) and the oracle."""
    from symboltz import surrogate as sg
    from symboltz.stream import MixedStream
    fams = workloads.mixed_families()
    ks = sg.k_grid()
    st = MixedStream(fams, ks)
    try:
        rows, cap = st.multi.rows(), st.multi.capacity()
        assert rows == [4, 2, 2, 2, 1], rows
        per = -(-2 * cap // len(ks))   # requests per family: at least twice the capacity in modes
        reqs = [r for r in workloads.mixed_requests(5 * per, seed=7) if r[0] in (0, 1)]
        say(f"case: {len(reqs)} requests of the n = 38 and n = 54 families x {len(ks)} k (capacity {cap}), "
            f"one launch (several-modes bodies) vs one launch per family")
        lp1, _, ret1, tm1 = st.solve(reqs)
        lp2, _, ret2, tm2 = st.solve(reqs, mode="serial")
        say(f"one launch {tm1['solve_ms']:.1f} ms, per family serial {tm2['solve_ms']:.1f} ms")
        assert np.all(ret1 == 0) and np.all(ret2 == 0)
        err = float(np.max(np.abs(np.expm1(lp1 - lp2))))
        say(f"one launch against per family: max |P1/P2 - 1| = {err:.2e}")
        assert err < 1e-6
        for f in (0, 1):
            i = next(j for j, r in enumerate(reqs) if r[0] == f)
            S = Setup(fams[f], reqs[i][1])
            P, r = S.O.spectrum_matter(k_dimensionless(ks[::32], reqs[i][1]["h"]))
            assert np.all(r["retcode"] == 0)
            e = np.max(np.abs(np.exp(lp1[i, ::32]) / P - 1))
            say(f"family {f}: max |P/P_oracle - 1| = {e:.2e}")
            assert e < 1e-4
    finally:
        st.close()


def test_stream_routes_requests_between_surrogate_and_solver():
    from symboltz import surrogate as sg
    from symboltz.stream import MixedStream
    fams = workloads.mixed_families()
    sf = workloads.MIXED_SURROGATE_FAMILY
    assert fams[sf] == sg.surrogate_model()
    ks = np.logspace(-4, np.log10(5.0), 32)
    say("case: surrogate (2 x 128 hidden) on 256 sets of the n = 54 family, stream of 20 requests")
    X = sg.lhs(256, seed=21)
    Y, _, _ = sg.solve_spectra(X, ks_mpc=ks)
    S = sg.Surrogate(n_members=2, hidden=(128, 128), n_out=len(ks), max_batch=4096)
    S.fit(X, Y, epochs=400, batch=64, lr=2e-3)
    reqs = workloads.mixed_requests(20, seed=3, out_of_box=0.5)
    inbox = np.array([r[2] is not None and bool(np.all((r[2] >= sg.LO) & (r[2] <= sg.HI))) for r in reqs])
    assert inbox.any() and np.any([r[0] == sf and not inbox[i] for i, r in enumerate(reqs)])
    st = MixedStream(fams, ks, surrogate=S, surrogate_family=sf, tol=1.0)
    plain = MixedStream(fams, ks)
    try:
        lp, ans, ret, tm = st.solve(reqs)
        lp0, ans0, ret0, _ = plain.solve(reqs)
        say(f"answered by the surrogate: {int(ans.sum())} of {len(reqs)}; timings {tm}")
        np.testing.assert_array_equal(ans, inbox)      # in-box requests of its family, nothing else
        assert np.all(ret == 0) and np.all(ret0 == 0) and not ans0.any()
        np.testing.assert_array_equal(lp[~ans], lp0[~ans])   # the solver's answers are the plain ones
        err = np.abs(np.exp(lp[ans] - lp0[ans]) - 1).max()
        say(f"surrogate answers against the solver: max |P/P_solver - 1| = {err:.2e}")
        assert err < 0.05
        # a tolerance no member pair meets sends every request to the solver
        st.tol = 0.0
        lp2, ans2, _, _ = st.solve(reqs)
        assert not ans2.any()
        np.testing.assert_array_equal(lp2, lp0)
    finally:
        st.close()
        plain.close()
        S.close()


def test_spectrum_matter_problems_one_launch_matches_per_problem():
    """The problem-level entry: problems of three models in one launch give the P(k) of
    `spectrum_matter` on each problem (the same arithmetic per mode)."""
    from symboltz import models
    from symboltz.problem import CosmologyProblem, remake, spectrum_matter, spectrum_matter_problems
    fams = [models.LCDM(lmax=10, h=False), models.w0waCDM(lmax=6), models.LCDM(lmax=6)]
    probs = []
    for M in fams:
        base = CosmologyProblem(M, models.parameters_Planck18(M))
        probs += [base, remake(base, {"h": 0.7})]
    ks = np.logspace(0, 3, 8)
    say(f"case: {len(probs)} problems over {len(fams)} models, {len(ks)} k, one launch")
    P = spectrum_matter_problems(probs, ks)
    for i, q in enumerate(probs):
        Pi = spectrum_matter(q, ks)
        Pi = Pi[0] if isinstance(Pi, tuple) else Pi
        err = float(np.max(np.abs(P[i] / np.asarray(Pi) - 1)))
        assert err < 1e-12, (i, err)
