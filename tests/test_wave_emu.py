"""The wavefront-per-mode kernel (sbz_wave_kernel) run on the CPU through the lane emulator
(tests/emu) against the oracle: final states, saved states and CMB sources ≤ 1e-4 relative, and
the failure return codes (k = 0, k = NaN give InitialFailure)."""
import numpy as np
import pytest

from conftest import Setup
from emu.emu import WaveEmu
from symboltz import models
from symboltz.solver import model_source


def rel(a, b):
    sc = np.maximum(np.abs(b).max(axis=-1, keepdims=True), 1e-300)
    return np.max(np.abs(a - b) / sc)


def test_wave_kernel_emulated_matches_oracle(fixture5):
    S = fixture5
    E = WaveEmu(model_source(S.M).full_src)
    ks = np.array([0.3, 30.0, 0.0, np.nan])
    t0 = np.clip(np.minimum(1e-2 / np.where(ks > 0, ks, 1.0), 1e-4), S.t[0], S.t[-1])
    saveat = np.linspace(0.05, S.t[-1], 9)
    g = E.solve(S.P, S.t, S.u, S.du, ks, t0, saveat=saveat, sources=True)
    o = S.O.solve(ks[:2], t0[:2], saveat=saveat)
    assert list(g["retcode"]) == [0, 0, 4, 4]
    assert rel(g["u"][:2], o["u"]) < 1e-4
    for i in range(2):
        assert rel(g["usave"][i], o["usave"][i]) < 1e-4
        ref = np.array([S.O.sources(ks[i], tau, o["usave"][i, j]) for j, tau in enumerate(saveat)])
        for c in range(3):
            assert np.max(np.abs(g["src"][i, :, c] - ref[:, c])) <= 1e-4 * (np.abs(ref[:, c]).max() + 1e-300)
    assert np.all(np.abs(g["stats"][:2, 0] - o["stats"][:, 0]) <= 0.05 * o["stats"][:, 0])


def test_wave_kernel_emulated_source_background_table(fixture5):
    """Sources read from the per-set background table (sbz_srcbg_kernel's rows) equal those of the
    per-mode evaluation; the states are untouched.  The save grid starts on a background knot and
    ends at τ0, the two places where the knot search could pick another interval."""
    S = fixture5
    E = WaveEmu(model_source(S.M).full_src)
    ks = np.array([0.05, 3.0, 60.0])
    t0 = np.clip(np.minimum(1e-2 / ks, 1e-4), S.t[0], S.t[-1])
    i0 = int(np.searchsorted(S.t, 1.0))
    saveat = np.concatenate([[S.t[i0]], np.linspace(S.t[i0], S.t[-1], 40)[1:]])
    assert saveat[-1] == S.t[-1]
    g0 = E.solve(S.P, S.t, S.u, S.du, ks, t0, saveat=saveat, sources=True)
    g1 = E.solve(S.P, S.t, S.u, S.du, ks, t0, saveat=saveat, sources=True, srcbg=True)
    assert np.all(g0["retcode"] == 0) and np.all(g1["retcode"] == 0)
    np.testing.assert_array_equal(g1["u"], g0["u"])
    np.testing.assert_array_equal(g1["usave"], g0["usave"])
    assert np.all(np.isfinite(g1["src"]))
    for c in range(3):
        sc = np.abs(g0["src"][..., c]).max() + 1e-300
        assert np.max(np.abs(g1["src"][..., c] - g0["src"][..., c])) <= 1e-13 * sc


def test_wave_kernel_emulated_both_occupancy_builds(fixture5):
    """sbz_wave_kernel forms the stage combinations in one pass over the stage vectors,
    sbz_wave_kernel_o2 in two; both match the oracle, and each other to rounding."""
    S = fixture5
    E = WaveEmu(model_source(S.M).full_src)
    ks = np.array([0.02, 2.0, 90.0])
    t0 = np.clip(np.minimum(1e-2 / ks, 1e-4), S.t[0], S.t[-1])
    g1 = E.solve(S.P, S.t, S.u, S.du, ks, t0)
    g2 = E.solve(S.P, S.t, S.u, S.du, ks, t0, o2=True)
    o = S.O.solve(ks, t0)
    assert np.all(g1["retcode"] == 0) and np.all(g2["retcode"] == 0)
    assert rel(g1["u"], o["u"]) < 1e-4 and rel(g2["u"], o["u"]) < 1e-4
    assert rel(g1["u"], g2["u"]) < 1e-8


def test_wave_kernel_emulated_factor_layouts_agree(fixture5):
    """sbz_wave_kernel_altlu keeps the border factors in the other LDS layout (split rows with
    unmasked solves, or one combined row per lane with masked solves): the same factor values meet
    the same operations, so the states and statistics agree bit for bit."""
    S = fixture5
    E = WaveEmu(model_source(S.M).full_src)
    ks = np.array([0.02, 2.0, 90.0])
    t0 = np.clip(np.minimum(1e-2 / ks, 1e-4), S.t[0], S.t[-1])
    g1 = E.solve(S.P, S.t, S.u, S.du, ks, t0)
    g2 = E.solve(S.P, S.t, S.u, S.du, ks, t0, altlu=True)
    assert np.all(g1["retcode"] == 0)
    np.testing.assert_array_equal(g2["u"], g1["u"])
    np.testing.assert_array_equal(g2["stats"], g1["stats"])


def test_wave_kernel_emulated_massive_nu_lmax16():
    M = models.LCDM(lmax_g=16, lmax_nu=16, lmax_h=10, nx=4)
    S = Setup(M, models.parameters_Planck18(M))
    E = WaveEmu(model_source(M).full_src)
    ks = np.array([0.5])
    t0 = np.clip(np.minimum(1e-2 / ks, 1e-4), S.t[0], S.t[-1])
    g = E.solve(S.P, S.t, S.u, S.du, ks, t0)
    o = S.O.solve(ks, t0)
    assert g["retcode"][0] == 0
    assert rel(g["u"], o["u"]) < 1e-4


def test_wave_kernel_emulated_w0wa():
    """The w0wa dark-energy fluid model (dark_energy.jl:43-64) through the wave kernel."""
    M = models.w0waCDM(lmax=6)
    S = Setup(M, dict(models.parameters_Planck18(M), w0=-0.9, wa=0.1))
    E = WaveEmu(model_source(M).full_src)
    ks = np.array([0.8, 40.0])
    t0 = np.clip(np.minimum(1e-2 / ks, 1e-4), S.t[0], S.t[-1])
    g = E.solve(S.P, S.t, S.u, S.du, ks, t0)
    o = S.O.solve(ks, t0)
    assert np.all(g["retcode"] == 0) and np.all(o["retcode"] == 0)
    assert rel(g["u"], o["u"]) < 1e-4


def test_wave4_kernel_emulated_matches_oracle_and_wave(massless10):
    """sbz_wave4_kernel (four modes per wavefront, 16 lanes each: DPP row broadcasts, row
    reductions, per-row step control) on the massless-ν lmax-10 model: 9 instances through the 4
    rows (rows refill from the queue and run out of work at different times), k = 0 and k = NaN
    fail with InitialFailure; states against the oracle (1e-4) and the one-mode-per-wavefront
    kernel (1e-8: the same arithmetic but for the summation order of the error norm)."""
    S = massless10
    E = WaveEmu(model_source(S.M).full_src)
    assert E.L.emu_has_wave4() == 1
    ks = np.array([100.0, 0.3, 0.0, 30.0, 1.0, np.nan, 50.0, 0.02, 10.0])
    kk = np.where(ks > 0, ks, 1.0)
    t0 = np.clip(np.minimum(1e-2 / kk, 1e-4), S.t[0], S.t[-1])
    g4 = E.solve(S.P, S.t, S.u, S.du, ks, t0, wave4=True)
    g1 = E.solve(S.P, S.t, S.u, S.du, ks, t0)
    good = np.isfinite(ks) & (ks > 0)
    assert list(g4["retcode"][~good]) == [4, 4]
    assert np.all(g4["retcode"][good] == 0)
    o = S.O.solve(ks[good], t0[good])
    assert rel(g4["u"][good], o["u"]) < 1e-4
    assert rel(g4["u"][good], g1["u"][good]) < 1e-8
    assert np.all(np.abs(g4["stats"][good, 0] - g1["stats"][good, 0]) <= 2)


def test_wave2h_kernel_emulated_matches_oracle_and_wave():
    """sbz_wave2h_kernel (two modes per wavefront, 32 lanes each: readlane broadcasts per half,
    row reductions across the half's two 16-lane rows, per-row step control) on the three-massive-ν
    lmax-6 model (border 26, chains 7: too wide for four modes per wavefront): 5 instances through
    the 2 rows, k = 0 and k = NaN fail with InitialFailure; states against the oracle (1e-4) and the
    one-mode-per-wavefront kernel (1e-8: the same arithmetic but for the summation order)."""
    M = models.LCDM(lmax=6)
    S = Setup(M, models.parameters_Planck18(M))
    E = WaveEmu(model_source(M).full_src)
    assert E.L.emu_has_wave4() == 0 and E.L.emu_has_wave2h() == 1
    ks = np.array([20.0, 0.0, 0.3, np.nan, 3.0])
    kk = np.where(ks > 0, ks, 1.0)
    t0 = np.clip(np.minimum(1e-2 / kk, 1e-4), S.t[0], S.t[-1])
    g2 = E.solve(S.P, S.t, S.u, S.du, ks, t0, wave2h=True)
    g1 = E.solve(S.P, S.t, S.u, S.du, ks, t0)
    good = np.isfinite(ks) & (ks > 0)
    assert list(g2["retcode"][~good]) == [4, 4]
    assert np.all(g2["retcode"][good] == 0)
    o = S.O.solve(ks[good], t0[good])
    assert rel(g2["u"][good], o["u"]) < 1e-4
    assert rel(g2["u"][good], g1["u"][good]) < 1e-8
    assert np.all(np.abs(g2["stats"][good, 0] - g1["stats"][good, 0]) <= 2)


def test_tangent_kernel_emulated_matches_oracle_differences(fixture5):
    """sbz_wave_tan_kernel (u and s = ∂u/∂log θ on one step sequence, the background's θ-dependence
    from two tight background solves at log θ ± h): u matches the plain kernel, and the total
    tangent of the final state, s + f(τ0) dτ0/dlog θ, matches the central difference of oracle
    solves at log θ ± h (rtol 1e-9) for θ = Ω_c0, one of the reference's AD test parameters
    (test/runtests.jl:363-377)."""
    from oracle import oracle
    from symboltz import _lib
    S = fixture5
    M = S.M
    E = WaveEmu(model_source(M).full_src)
    assert E.L.emu_has_tan() == 1
    ks = np.array([1.0, 30.0])
    t0 = np.clip(np.minimum(1e-2 / ks, 1e-4), S.t[0], S.t[-1])
    h = 1e-3
    g1 = E.solve(S.P, S.t, S.u, S.du, ks, t0)
    for name in ("Omega_c0",):   # Ω_b0 as well on the GPU (test_configs.py)
        x = np.log(S.pars[name])
        tight = []
        for sgn in (1, -1):
            P = models.make_params(M, dict(S.pars, **{name: np.exp(x + sgn * h)}))
            tight.append(_lib.background_solve(P, 1e-10, 1e-10, 100000)[:4])
        g = E.solve_tangent([(S.P, S.t, S.u, S.du)] + tight, ks, t0, 0.5 / h)
        assert np.all(g["retcode"] == 0)
        assert rel(g["u"], g1["u"]) < 1e-8
        assert np.all(g["stats"][:, 0] == g1["stats"][:, 0])
        dtau0 = (tight[0][1][-1] - tight[1][1][-1]) / (2 * h)
        s_tot = g["s"] + g["fend"] * dtau0
        fd = []
        for P, t, u, du in tight:
            O = oracle.Oracle(M, P, t, u, du)
            o = O.solve(ks, t0, rtol=1e-9, atol=1e-9, maxiters=10**6)
            assert np.all(o["retcode"] == 0)
            fd.append(o["u"])
        fd = (fd[0] - fd[1]) / (2 * h)
        err = np.max(np.abs(s_tot - fd) / np.abs(fd).max(axis=1, keepdims=True), axis=1)
        print(name, err)
        assert np.all(err < 1e-3), (name, err)
