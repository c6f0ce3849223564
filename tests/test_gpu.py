"""GPU parity: the HIP path through the C ABI against the oracle on the same inputs.

The plain GPU path must match the CPU restatement of the same algorithm to ≤ 1e-4 relative in
P(k) and in the perturbation variables (final states, compared relative to each mode's largest
component); it is also held to the reference's CLASS fixture (norm-wise 1e-3,
test/runtests.jl:872-876) and to its failure semantics (k = 0, k = NaN fail without raising,
test/runtests.jl:31,360)."""
import sys
import warnings

import numpy as np
import pytest

from conftest import GOLDEN, Setup
from symboltz import models
from symboltz.problem import solve, solve_sets, spectrum_matter, CosmologyProblem, issuccess
from symboltz.solver import GpuModel

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def gpu_solve(S, ks, t0, **kw):
    gm = GpuModel(S.M)
    gm.upload_background(0, S.t, S.u, S.du, S.P)
    plan = gm.plan(np.zeros(len(ks), dtype=np.int32), ks, t0, **kw)
    ms = plan.run()
    r = plan.fetch()
    r["pk"] = plan.pk()
    r["ms"] = ms
    plan.close()
    return r


def rel_state_err(a, b):
    scale = np.maximum(np.abs(b).max(axis=1, keepdims=True), 1e-300)
    return np.max(np.abs(a - b) / scale)


def test_fixture_model_matches_oracle(fixture5):
    S = fixture5
    say(f"case: ΛCDM lmax=5, 24 modes, n={S.O.n}")
    ks = 10 ** np.linspace(np.log10(0.05), np.log10(5000.0), 24)
    t0 = np.clip(np.minimum(1e-2 / ks, 1e-4), S.t[0], S.t[-1])
    g = gpu_solve(S, ks, t0)
    o = S.O.solve(ks, t0)
    assert np.all(g["retcode"] == 0)
    assert np.all(o["retcode"] == 0)
    Po = np.array([S.O.p0(k) * S.O.delta_m(k, S.t[-1], o["u"][i]) ** 2 for i, k in enumerate(ks)])
    assert np.max(np.abs(g["pk"] / Po - 1)) < 1e-4
    assert rel_state_err(g["u"], o["u"]) < 1e-4
    # identical algorithm: step counts agree closely
    assert np.max(np.abs(g["stats"][:, 0] - o["stats"][:, 0]) / o["stats"][:, 0]) < 0.05


def test_class_pk_fixture_on_gpu(fixture5):
    S = fixture5
    say("case: CLASS P(k) fixture, 516 modes")
    d = np.loadtxt(GOLDEN / "class_Pk.dat")
    ks, Pc = d[:, 0], d[:, 1]
    t0 = np.clip(np.minimum(1e-2 / ks, 1e-4), S.t[0], S.t[-1])
    g = gpu_solve(S, ks, t0)
    assert np.all(g["retcode"] == 0)
    P = g["pk"]
    assert np.linalg.norm(P - Pc) / max(np.linalg.norm(P), np.linalg.norm(Pc)) < 1e-3


def test_massless_lmax10_matches_oracle(massless10):
    S = massless10
    say("case: ΛCDM massless ν lmax=10, 40 modes in [1e-4, 1]/Mpc")
    h = S.pars["h"]
    from symboltz.constants import k_dimensionless
    ks = k_dimensionless(10 ** np.linspace(-4, 0, 40), h)
    t0 = np.clip(np.minimum(1e-2 / ks, 1e-4), S.t[0], S.t[-1])
    g = gpu_solve(S, ks, t0)
    o = S.O.solve(ks, t0)
    Po = np.array([S.O.p0(k) * S.O.delta_m(k, S.t[-1], o["u"][i]) ** 2 for i, k in enumerate(ks)])
    assert np.all(g["retcode"] == 0)
    assert np.max(np.abs(g["pk"] / Po - 1)) < 1e-4
    assert rel_state_err(g["u"], o["u"]) < 1e-4


def test_failures_are_reported_not_raised(fixture5):
    say("case: failure semantics k=0, k=NaN")
    M = models.LCDM(lmax=5)
    prob = CosmologyProblem(M, models.parameters_Planck18(M))
    with warnings.catch_warnings(record=True) as w:
        warnings.simplefilter("always")
        sol = solve(prob, np.array([1.0, 0.0, np.nan]))
    assert sol.pts["retcode"][0] == 0
    assert sol.pts["retcode"][1] != 0 and sol.pts["retcode"][2] != 0
    assert not issuccess(sol)
    assert any("Perturbation (mode k = nan) solution failed" in str(x.message) for x in w)
    with pytest.raises(TypeError):
        solve(prob, np.array([1.0 + 1.0j]))


def test_saveat_dense_output_matches_oracle(fixture5):
    S = fixture5
    say("case: saveat 50 times, 6 modes")
    ks = np.array([0.3, 3.0, 30.0, 300.0, 1000.0, 3000.0])
    t0 = np.full(len(ks), S.t[0])
    saveat = np.linspace(0.01, S.t[-1], 50)
    gm = GpuModel(S.M)
    gm.upload_background(0, S.t, S.u, S.du, S.P)
    plan = gm.plan(np.zeros(len(ks), dtype=np.int32), ks, t0, saveat=saveat)
    plan.run()
    g = plan.fetch(save=True)
    plan.close()
    o = S.O.solve(ks, t0, saveat=saveat)
    for i in range(len(ks)):
        assert rel_state_err(g["usave"][i], o["usave"][i]) < 1e-4


def test_batched_parameter_sets_match_oracle():
    say("case: 4 parameter sets x 12 modes in one launch")
    M = models.LCDM(lmax=5)
    base = models.parameters_Planck18(M)
    sets = []
    for i, (dh, dc) in enumerate([(0.0, 0.0), (0.03, 0.01), (-0.04, -0.015), (0.05, 0.02)]):
        p = dict(base)
        p["h"] = base["h"] + dh
        p["Omega_c0"] = (0.12 + dc) / p["h"] ** 2
        p["Omega_b0"] = 0.0224 / p["h"] ** 2
        sets.append(p)
    ks = 10 ** np.linspace(-1, 3, 12)
    pk, res, bgs = solve_sets(M, sets, ks)
    assert np.all(res["retcode"] == 0)
    for s, pars in enumerate(sets):
        S = Setup(M, pars)
        P, r = S.O.spectrum_matter(ks)
        assert np.max(np.abs(pk[s] / P - 1)) < 1e-4


def test_cmb_sources_match_oracle(fixture5):
    S = fixture5
    say("case: CMB sources on 60 times, 8 modes")
    from symboltz.solver import SAVE_SOURCES, SAVE_STATES
    ks = np.array([0.5, 2.0, 8.0, 30.0, 100.0, 300.0, 900.0, 2500.0])
    t0 = np.full(len(ks), S.t[0])
    saveat = np.linspace(0.02, S.t[-1], 60)
    gm = GpuModel(S.M)
    gm.upload_background(0, S.t, S.u, S.du, S.P)
    plan = gm.plan(np.zeros(len(ks), dtype=np.int32), ks, t0, saveat=saveat, save=SAVE_SOURCES | SAVE_STATES)
    plan.run()
    g = plan.fetch(save=True)
    src = plan.sources()
    plan.close()
    o = S.O.solve(ks, t0, saveat=saveat)
    for i, k in enumerate(ks):
        ref = np.array([S.O.sources(k, tau, o["usave"][i, j]) for j, tau in enumerate(saveat)])
        for c in range(3):
            scale = np.abs(ref[:, c]).max() + 1e-300
            assert np.max(np.abs(src[i, :, c] - ref[:, c])) <= 1e-4 * scale, (k, c)


def test_w0wa_model_matches_oracle():
    say("case: w0waCDM lmax 6 (w0 = -0.9, wa = 0.1), 10 modes")
    M = models.w0waCDM(lmax=6)
    S = Setup(M, dict(models.parameters_Planck18(M), w0=-0.9, wa=0.1))
    ks = 10 ** np.linspace(np.log10(0.1), np.log10(3000.0), 10)
    t0 = np.clip(np.minimum(1e-2 / ks, 1e-4), S.t[0], S.t[-1])
    g = gpu_solve(S, ks, t0)
    o = S.O.solve(ks, t0)
    assert np.all(g["retcode"] == 0) and np.all(o["retcode"] == 0)
    Po = np.array([S.O.p0(k) * S.O.delta_m(k, S.t[-1], o["u"][i]) ** 2 for i, k in enumerate(ks)])
    assert np.max(np.abs(g["pk"] / Po - 1)) < 1e-4
    assert rel_state_err(g["u"], o["u"]) < 1e-4


def test_gpu_background_batch_matches_host_solver():
    # the batched GPU background solve (one thread per set) against the host solver, set by set
    from scipy.stats import qmc
    from symboltz import _lib
    say("case: GPU background solve, 24 parameter sets (one failing)")
    M = models.LCDM(lmax=5)
    base = models.parameters_Planck18(M)
    X = qmc.scale(qmc.LatinHypercube(d=5, seed=3).random(24), [0.60, 0.020, 0.10, 2.9, 0.92],
                  [0.75, 0.024, 0.14, 3.2, 1.00])
    sets = [dict(base, h=h, Omega_b0=wb / h**2, Omega_c0=wc / h**2, ln_As1e10=la, ns=ns) for h, wb, wc, la, ns in X]
    sets[7]["Omega_c0"] = float("nan")      # fails (the reference warns "Background solution failed")
    Ps = np.stack([models.make_params(M, p) for p in sets])
    P, knots, nrej, status, ms = _lib.background_solve_batch(Ps)
    say(f"batch kernel {ms:.1f} ms")
    assert status[7] != 0 and np.all(np.delete(status, 7) == 0)
    for i in range(24):
        if i == 7:
            continue
        Ph, t, u, du, _ = _lib.background_solve(Ps[i])
        tg, ug, dug = knots[i]
        for slot in (models.P_TAU0, models.P_KAPPA0):
            assert abs(P[i, slot] / Ph[slot] - 1) < 1e-8, (i, slot)
        # τrec is the visibility maximum over the knots (solve.jl:181-195): equal up to the knot
        # spacing there, which differs with the rounding of the step sequence
        j = np.searchsorted(t, Ph[models.P_TAUREC])
        assert abs(P[i, models.P_TAUREC] - Ph[models.P_TAUREC]) <= 2 * (t[j + 1] - t[j - 1]), i
        # the step sequences differ by rounding (device and host math differ in the last bits), so
        # the knots differ; the two Hermite splines (the background the perturbations read) agree
        # at sampled times to the integration tolerance
        ts = np.geomspace(t[0] * 10, t[-1] * 0.999, 60)
        errs = []
        for c in (0, 4):   # a and _κ carry exact knot derivatives
            a = hermite(tg, ug[:, c], dug[:, c], ts)
            b = hermite(t, u[:, c], du[:, c], ts)
            errs.append(np.max(np.abs(a - b) / (np.abs(b) + 1e-3 * np.abs(b).max())))
        say(f"set {i}: knots {len(tg)} (GPU) / {len(t)} (host); spline differences a {errs[0]:.1e}, κ {errs[1]:.1e}")
        assert max(errs) <= 1e-5, i


def hermite(t, y, dy, x):
    i = np.clip(np.searchsorted(t, x) - 1, 0, len(t) - 2)
    h = t[i + 1] - t[i]
    s = (x - t[i]) / h
    return ((1 + 2 * s) * (1 - s) ** 2 * y[i] + s * (1 - s) ** 2 * h * dy[i] + s * s * (3 - 2 * s) * y[i + 1]
            + s * s * (s - 1) * h * dy[i + 1])


def test_spectrum_matter_jacobian():
    # sensitivities by central differences with every shifted cosmology in one launch; pinned by
    # the closed forms d log P / d log A_s = 1 and d log P / d log n_s = n_s log(k / k_pivot)
    # (P0 ∝ A_s (k/k_p)^(n_s - 1), the transfer function depends on neither), and for Ω_c0, Ω_b0
    # against the oracle's differences at the reference's tolerance (test/runtests.jl:363-377)
    from symboltz.problem import spectrum_matter_jacobian
    from symboltz.models import P_KPIVOT
    say("case: d log P / d log θ for As, ns, Omega_c0, Omega_b0 at 20 k (9 cosmologies, one launch)")
    M = models.LCDM(lmax=5)
    pars = dict(models.parameters_Planck18(M))
    pars.pop("As", None)
    prob = CosmologyProblem(M, pars)
    ks = 10 ** np.linspace(0, 3, 20)
    J, P = spectrum_matter_jacobian(prob, ks, ["ln_As1e10", "ns", "Omega_c0", "Omega_b0"])
    lnAs = pars["ln_As1e10"]
    # ln_As1e10 = log(1e10 A_s): d log P / d log(ln_As1e10) = ln_As1e10 exactly
    assert np.max(np.abs(J[:, 0] - lnAs)) < 5e-6
    kp = prob.P[P_KPIVOT]
    assert np.max(np.abs(J[:, 1] - pars["ns"] * np.log(ks / kp))) < 5e-6
    S = Setup(M, pars)
    base = dict(pars)
    for j, n in ((2, "Omega_c0"), (3, "Omega_b0")):
        x = np.log(base[n])
        h = 1e-3 * max(abs(x), 1.0)
        Pp = Setup(M, dict(base, **{n: np.exp(x + h)})).O.spectrum_matter(ks)[0]
        Pm = Setup(M, dict(base, **{n: np.exp(x - h)})).O.spectrum_matter(ks)[0]
        Jo = (np.log(Pp) - np.log(Pm)) / (2 * h)
        assert np.max(np.abs(J[:, j] - Jo)) < 1e-3, n


def test_massive_nu_n100_with_sources_matches_oracle():
    # the benchmark's model (photon and massless-ν lmax 16, one massive ν with 4 bins and lmax 10,
    # n = 100): P(k), final states and CMB sources against the oracle
    from symboltz.solver import SAVE_SOURCES, SAVE_STATES
    say("case: n = 100 massive-ν model, 12 modes in [1e-4, 0.5]/Mpc, sources on 80 τ")
    M = models.LCDM(lmax_g=16, lmax_nu=16, lmax_h=10, nx=4)
    pars = dict(models.parameters_Planck18(M), Neff=2.046, N=1, m_eV=0.06)
    S = Setup(M, pars)
    from symboltz.constants import k_dimensionless
    ks = k_dimensionless(10 ** np.linspace(-4, np.log10(0.5), 12), pars["h"])   # below the ill-conditioned ISW range
    t0 = np.clip(np.minimum(1e-2 / ks, 1e-4), S.t[0], S.t[-1])
    saveat = np.linspace(max(S.t[0], 1e-2), S.t[-1], 80)
    gm = GpuModel(S.M)
    gm.upload_background(0, S.t, S.u, S.du, S.P)
    plan = gm.plan(np.zeros(len(ks), dtype=np.int32), ks, t0, saveat=saveat, save=SAVE_SOURCES | SAVE_STATES)
    plan.run()
    g = plan.fetch(save=True)
    src = plan.sources()
    pk = plan.pk()
    plan.close()
    o = S.O.solve(ks, t0, saveat=saveat, sources=True)
    assert np.all(g["retcode"] == 0) and np.all(o["retcode"] == 0)
    Po = np.array([S.O.p0(k) * S.O.delta_m(k, S.t[-1], o["u"][i]) ** 2 for i, k in enumerate(ks)])
    assert np.max(np.abs(pk / Po - 1)) < 1e-4
    assert rel_state_err(g["u"], o["u"]) < 1e-4
    for i in range(len(ks)):
        for c in range(3):
            scale = np.abs(o["src"][i, :, c]).max() + 1e-300
            assert np.max(np.abs(src[i, :, c] - o["src"][i, :, c])) <= 1e-4 * scale, (i, c)


def test_high_lmax_32_matches_oracle():
    # the reference's "High lmax" test (test/runtests.jl:808-812): ΛCDM(lmax = 32), n = 236, solved
    # from the background start at k = 0.1 .. 1000; here also against the oracle
    say("case: ΛCDM lmax 32 (n = 236), 5 modes from the background start")
    M = models.LCDM(lmax=32)
    prob = CosmologyProblem(M, models.parameters_Planck18(M))
    ks = np.array([1e-1, 1e0, 1e1, 1e2, 1e3])
    sol = solve(prob, ks)
    assert issuccess(sol)
    S = Setup(M)
    o = S.O.solve(ks, np.full(len(ks), S.t[0]))
    assert np.all(o["retcode"] == 0)
    assert rel_state_err(sol.pts["u"], o["u"]) < 1e-4
    Po = np.array([S.O.p0(k) * S.O.delta_m(k, S.t[-1], o["u"][i]) ** 2 for i, k in enumerate(ks)])
    assert np.max(np.abs(sol.pts["pk"] / Po - 1)) < 1e-4


def test_spectrum_matter_modes_and_times(fixture5):
    # spectrum_matter(modes, prob, k, τ) (test/runtests.jl:609-622): shapes, "m" equal to the total
    # matter spectrum, "cbh" equal to "m", and the CDM and baryon spectra at τ = 1.5, 3.0 against
    # Δ_s = δ_s + 3ℋθ_s/k² from the oracle's saved states
    from symboltz.problem import spectrum_matter_modes
    say("case: P(k, τ) for modes m, c, b, cb, cbh, h at 2 times, 4 k")
    S = fixture5
    prob = CosmologyProblem(S.M, S.pars)
    ks = np.array([1e-4, 1e-3, 1e-2, 1e-1])
    taus = np.array([1.5, 3.0])
    modes = ["m", "c", "b", "cb", "cbh", "h"]
    P = spectrum_matter_modes(modes, prob, ks, taus, k_unit="1/Mpc")
    assert P.shape == (6, 2, 4)
    P0 = spectrum_matter_modes(modes, prob, ks, k_unit="1/Mpc")
    assert P0.shape == (6, 4)
    Pm, _ = spectrum_matter(prob, ks, k_unit="1/Mpc")
    assert np.max(np.abs(P0[0] / Pm - 1)) < 1e-12
    assert np.array_equal(P0[4], P0[0]) and np.array_equal(P[4], P[0])
    from symboltz.constants import k_dimensionless
    kd = k_dimensionless(ks, S.pars["h"])
    t0 = np.clip(np.minimum(1e-2 / kd, 1e-4), S.t[0], S.t[-1])
    o = S.O.solve(kd, t0, saveat=taus)
    L = S.O.layout()
    for it, tau in enumerate(taus):
        H = S.O.background(tau)["H"]
        for j, col in ((1, L["C"]), (2, L["B"])):
            D = o["usave"][:, it, col] + 3 * H * o["usave"][:, it, col + 1] / kd**2
            Po = np.array([S.O.p0(k) for k in kd]) * D**2
            assert np.max(np.abs(P[j, it] / Po - 1)) < 1e-4, (modes[j], tau)
    with pytest.raises(ValueError):
        spectrum_matter_modes(["x"], prob, ks)


def test_power_spectrum_converged_to_0p1_percent():
    # test/runtests.jl:624-630: P(k) at the default tolerances within 0.1 % of a tight solve
    # (rtol = atol = 1e-10, every mode from the background start) for 100 k in [0.1, 1e4] H0/c
    say("case: P(k) rtol 1e-5 vs 1e-10, 100 k in [0.1, 1e4]")
    M = models.LCDM(lmax=5)
    prob = CosmologyProblem(M, models.parameters_Planck18(M))
    ks = 10 ** np.linspace(-1, 4, 100)
    P0, sol0 = spectrum_matter(prob, ks, ktini=0.0, tauinimax=0.0,
                               ptopts=dict(reltol=1e-10, abstol=1e-10, maxiters=10_000_000))
    P, _ = spectrum_matter(prob, ks)
    assert issuccess(sol0)
    err = np.abs(P / P0 - 1)
    say(f"max |P/P0 - 1| = {err.max():.2e}")
    assert np.all(err < 1e-3)


def test_stability_over_latin_hypercube():
    # test/runtests.jl:505-546: every parameter varied by ±50 % around Planck18, 100 Latin-hypercube
    # samples, k = 1, 10, 100, 1000 — all solves succeed (with and without reionization)
    from scipy.stats import qmc
    from symboltz.models import P_REION
    say("case: 2 x 100 Latin-hypercube cosmologies (every parameter ±50 %), 4 k, one launch each")
    M = models.LCDM(lmax=5)
    base = models.parameters_Planck18(M)
    names = list(base)
    lo = np.array([0.5 * base[n] for n in names])
    hi = np.array([1.5 * base[n] for n in names])
    X = qmc.scale(qmc.LatinHypercube(d=len(names), seed=0).random(100), lo, hi)
    sets = [dict(zip(names, x)) for x in X]
    ks = np.array([1e0, 1e1, 1e2, 1e3])
    pk, res, bgs = solve_sets(M, sets, ks)
    assert np.all(res["retcode"] == 0) and np.all(np.isfinite(pk)) and np.all(pk > 0)
    # reionization switched off
    from symboltz.problem import solvebg_sets
    probs = [CosmologyProblem(M, p) for p in sets]
    for p in probs:
        p.P[P_REION] = 0.0
    bgs = solvebg_sets(probs)
    gm = probs[0].gpu
    for si, bg in enumerate(bgs):
        gm.upload_background(si, bg.t, bg.u, bg.du, bg.P)
    n = len(ks)
    t0 = np.concatenate([np.clip(np.minimum(1e-2 / ks, 1e-4), bg.t[0], bg.t[-1]) for bg in bgs])
    plan = gm.plan(np.repeat(np.arange(len(bgs), dtype=np.int32), n), np.tile(ks, len(bgs)), t0)
    plan.run()
    r = plan.fetch()
    pk2 = plan.pk()
    plan.close()
    assert np.all(r["retcode"] == 0) and np.all(np.isfinite(pk2)) and np.all(pk2 > 0)


def test_solution_interpolation_and_accessing(fixture5):
    # sol(is, τs, ks) (src/solve.jl:720-787; test/runtests.jl:35-66,89-96): output shapes follow the
    # arguments, wavenumbers outside the solved range raise, a solved wavenumber returns its own
    # mode, values between modes are linear in log k, and the background reads the knot spline
    S = fixture5
    say("case: solution interpolation, 2 solved modes + 20 modes in [1e-2, 1e1]")
    M = S.M
    prob = CosmologyProblem(M, S.pars)
    ks = np.array([1e2, 1e3])
    taus = np.array([1.0, 2.0, 3.0])
    sol = solve(prob, ks)
    iPhi = 0
    ii = [iPhi, iPhi, 1, 2]
    assert np.ndim(sol(iPhi, taus[0], ks[0])) == 0
    assert np.shape(sol(ii, taus[0], ks[0])) == (4,)
    assert np.shape(sol(iPhi, taus, ks[0])) == (3,)
    assert np.shape(sol(ii, taus, ks[0])) == (4, 3)
    assert np.shape(sol(iPhi, taus[0], ks)) == (2,)
    assert np.shape(sol(ii, taus[0], ks)) == (4, 2)
    assert np.shape(sol(iPhi, taus, ks)) == (3, 2)
    assert np.shape(sol(ii, taus, ks)) == (4, 3, 2)
    with pytest.raises(ValueError, match="below the minimum solved wavenumber"):
        sol(ii, taus, ks[0] - 1)
    with pytest.raises(ValueError, match="above the maximum solved wavenumber"):
        sol(ii, taus, ks[-1] + 1)
    assert np.shape(sol("a", taus)) == (3,) and abs(sol("a", S.t[-1]) - 1.0) < 1e-12
    with pytest.raises(ValueError, match="No perturbations"):
        solve(prob)(ii, taus, ks)
    # values: the states at τ0 equal the solve's final states; between modes, log-k linear
    ks2 = 10 ** np.linspace(-2, 1, 20)
    sol2 = solve(prob, ks2)
    u0 = sol2(list(range(6)), S.t[-1], ks2)           # [6][n_k]
    assert np.max(np.abs(u0.T - sol2.pts["u"][:, :6])) <= 1e-12 * np.abs(sol2.pts["u"][:, :6]).max()
    km = np.sqrt(ks2[3] * ks2[4])
    v = sol2(list(range(6)), S.t[-1], km)
    assert np.allclose(v, 0.5 * (sol2.pts["u"][3, :6] + sol2.pts["u"][4, :6]), rtol=1e-12, atol=0)


def test_config4_model_families_match_oracle():
    # the mixed workload of four model families packed by state size (n = 38, 56, 100, 236): each
    # family's sets run in one launch; Latin-hypercube parameters over h, ω_b, ω_c, ln 1e10 As,
    # ns, and for the families that have them w0, wa and Σm_ν ∈ [0.06, 0.30] eV (3 degenerate
    # species); P(k) against the oracle.  w(a) = w0 + wa (1 - a) is kept on one side of -1: the
    # fluid equations divide by 1 + w (c_s²/(1+w), ẇ/(1+w), dark_energy.jl:55-57), so a set whose
    # w crosses -1 is singular there; such a set must fail per mode without raising (below).
    from scipy.stats import qmc
    say("case: 4 model families x 3 LHS sets x 6 k, P(k) against the oracle")
    families = [models.LCDM(lmax=10, h=False), models.w0waCDM(lmax=6),
                models.LCDM(lmax_g=16, lmax_nu=16, lmax_h=10, nx=4), models.LCDM(lmax=32)]
    ks = 10 ** np.linspace(-1, 3, 6)
    X = qmc.scale(qmc.LatinHypercube(d=8, seed=4).random(3), [0.60, 0.020, 0.10, 2.9, 0.92, -0.95, 0.0, 0.06],
                  [0.75, 0.024, 0.14, 3.2, 1.00, -0.8, 0.3, 0.30])
    for M in families:
        base = models.parameters_Planck18(M)
        sets = []
        for h, wb, wc, la, ns, w0, wa, mnu in X:
            p = dict(base, h=h, Omega_b0=wb / h**2, Omega_c0=wc / h**2, ln_As1e10=la, ns=ns)
            if "w0" in base:
                p.update(w0=w0, wa=wa)
            if "m_eV" in base:
                p.update(N=3, m_eV=mnu / 3) if base.get("N", 1) == 3 else p.update(m_eV=mnu)
            sets.append(p)
        pk, res, bgs = solve_sets(M, sets, ks)
        assert np.all(res["retcode"] == 0), M.name
        for si, p in enumerate(sets):
            P, r = Setup(M, p).O.spectrum_matter(ks)
            assert np.all(r["retcode"] == 0)
            assert np.max(np.abs(pk[si] / P - 1)) < 1e-4, (M.name, M.n_unknowns, si)
    # a w0wa set crossing w = -1 (w0 = -1.1, wa = 0.3: w = -0.8 early, -1.1 today)
    M = models.w0waCDM(lmax=6)
    cross = dict(models.parameters_Planck18(M), w0=-1.1, wa=0.3)
    with warnings.catch_warnings():
        warnings.simplefilter("ignore")
        pk, res, _ = solve_sets(M, [cross], ks)
    assert np.any(res["retcode"] != 0)
