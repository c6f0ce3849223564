"""CMB angular power spectra (src/observables/angular.jl:198-359; fourier.jl:419-532).

CPU: the oracle pipeline against the reference's CLASS fixture test/class_Cl.dat with the
reference's own gate (rtol 2e-3 norm-wise for TT, EE and ψψ, test/runtests.jl:879-886), the
Chebyshev interpolator against polynomial exactness, and the spline quadrature weights against
scipy's natural cubic spline integral.
GPU: the product pipeline (`symboltz.cmb.spectrum_cmb`) against the same fixture and against the
oracle pipeline; its device kernels against the oracle's on the same coarse sources."""
import numpy as np
import pytest
from scipy.interpolate import CubicSpline

from conftest import GOLDEN
from oracle import oracle_cmb as oc
from symboltz.cmb import ChebyshevInterpolator, cosgrid, fk_inv_tanh, fk_tanh, lingrid, spline_integral_weights
from symboltz.problem import CosmologyProblem, spectrum_primordial


def fixture_ls():
    d = np.loadtxt(GOLDEN / "class_Cl.dat")
    lsc = d[:, 0]
    ls = np.unique(np.round(np.exp(np.linspace(np.log(lsc[0]), np.log(lsc[-1]), 200))).astype(int))
    return d, lsc, ls


def normwise(a, b):
    # Julia's isapprox(a, b; rtol): norm(a - b) ≤ rtol max(norm(a), norm(b))
    return np.linalg.norm(a - b) / max(np.linalg.norm(a), np.linalg.norm(b))


def test_spline_integral_weights_match_natural_spline_integral():
    rng = np.random.default_rng(0)
    x = np.concatenate([[0.0], np.sort(rng.uniform(0.01, 100.0, 400))])
    y = rng.standard_normal((len(x), 3))
    ref = CubicSpline(x, y, bc_type="natural").integrate(x[0], x[-1])
    assert np.max(np.abs(spline_integral_weights(x) @ y - ref)) < 1e-12 * np.abs(y).sum()
    # exact for a straight line (natural spline reproduces it) on any grid
    assert abs(spline_integral_weights(x) @ (2 * x + 1) - (x[-1] ** 2 + x[-1])) < 1e-9


def test_chebyshev_interpolator_is_exact_on_polynomials():
    c = ChebyshevInterpolator(1e-2, 1e4, 30, f=fk_tanh, finv=fk_inv_tanh)
    assert c.xs[0] == 1e4 and c.xs[-1] == 1e-2 and np.all(np.diff(c.xs) < 0)
    rng = np.random.default_rng(1)
    coef = rng.standard_normal(31)
    p = np.polynomial.Polynomial(coef, domain=[c.ys[-1], c.ys[0]])
    for x in rng.uniform(1e-2, 1e4, 50):
        assert abs(c(p(c.ys), x) - p(float(fk_tanh(x)))) < 1e-9 * np.abs(coef).sum()
    # the oracle's interpolation (C++) agrees with the product's host evaluation
    ch = oc.Cheb(1e-2, 1e4, 30, f=lambda k: np.tanh(k / 2000.0), finv=lambda y: 2000.0 * np.arctanh(y))
    assert np.array_equal(ch.xs, c.xs) and np.array_equal(ch.ws, c.ws)
    xs = rng.uniform(1e-2, 1e4, 20)
    o = oc.kinterp(p(c.ys)[:, None], ch, xs)[:, 0]
    assert np.max(np.abs(o - np.array([c(p(c.ys), x) for x in xs]))) < 1e-12 * np.abs(coef).sum()
    # at a node the node value is returned
    assert oc.kinterp(p(c.ys)[:, None], ch, [c.xs[5]])[0, 0] == p(c.ys)[5]


def test_grids_follow_the_reference():
    g = lingrid(0.01, 1e4, step=np.pi / 3.3)
    assert g[0] == 0.01 and g[-1] == 1e4 and len(g) == int(np.ceil((1e4 - 0.01) / (np.pi / 3.3))) + 1
    c = cosgrid(0.0, 1.0, length=300)
    assert c[0] == 0.0 and c[-1] == 1.0 and np.all(np.diff(np.diff(c)) > 0)
    assert oc.cosgrid(0.0, 1.0, 300)[-1] == 1.0


def test_level_refinement_equals_queue_refinement():
    # the generation-by-generation bisection (all midpoints of a generation in one launch) picks
    # exactly the wavenumbers of the reference's interval queue, on an analytic "source"
    from symboltz.cmb import refine_levels
    taus = np.linspace(0.1, 3.0, 7)

    def S_of(k):
        k = np.atleast_1d(k)[:, None]
        t = taus[None, :]
        return np.stack([np.sin(k * t / 40.0) * np.exp(-k / 900.0), np.cos(k / 70.0) / (1 + t)], axis=2)

    for rtol, f, finv in ((3e-2, np.log, np.exp), (2e-2, lambda x: x, lambda y: y)):
        ks = np.geomspace(1.0, 3000.0, 9)
        its = slice(0, len(taus) - 1)
        K1, S1, calls = refine_levels(S_of, ks, f, finv, rtol, 0.0, its)
        K2, S2 = oc.refine_queue(lambda k: S_of(k)[0], ks, f, finv, rtol, 0.0, its)
        o = np.argsort(K1, kind="stable")
        assert len(K1) > 40 and np.array_equal(K1[o], K2) and np.array_equal(S1[o], S2)
        assert calls < len(K1) / 4


def test_oracle_cmb_spectra_match_class_fixture(fixture5):
    S = fixture5
    d, lsc, ls = fixture_ls()
    prob = CosmologyProblem(S.M, S.pars)
    D, det = oc.spectrum_cmb(S.O, lambda k: spectrum_primordial(k, prob), ["TT", "EE", "ψψ"], ls, ls_fine=lsc)
    for i, name in enumerate(["TT", "EE", "ψψ"]):
        err = normwise(D[:, i], d[:, 1 + i])
        print(f"oracle {name}: norm-wise {err:.2e}")
        assert err < 2e-3, name


@pytest.mark.gpu
def test_gpu_cmb_spectra_match_class_fixture_and_oracle(fixture5):
    from symboltz.bessel import SphericalBesselCache
    from symboltz.cmb import spectrum_cmb
    print("case: C_l TT/EE/ψψ for the CLASS fixture, 131 source modes, l ≤ 2500", flush=True)
    S = fixture5
    d, lsc, ls = fixture_ls()
    prob = CosmologyProblem(S.M, S.pars)
    jl = SphericalBesselCache(ls)
    det, tm = {}, {}
    D = spectrum_cmb(["TT", "EE", "ψψ"], prob, jl, lsc, normalization="Dl", timing=tm, details=det)
    print({k: v for k, v in tm.items() if k.endswith("_ms")}, flush=True)
    for i, name in enumerate(["TT", "EE", "ψψ"]):
        err = normwise(D[:, i], d[:, 1 + i])
        print(f"gpu {name}: norm-wise {err:.2e}")
        assert err < 2e-3, name
    # the same pipeline on the CPU oracle (its own source solve): the two differ only by the
    # integration of the sources (same algorithm)
    Do, dto = oc.spectrum_cmb(S.O, lambda k: spectrum_primordial(k, prob), ["TT", "EE", "ψψ"], ls, ls_fine=lsc)
    for i in range(3):
        assert normwise(D[:, i], Do[:, i]) < 1e-4
    # kernels alone: the oracle's interpolation and LOS on the GPU's own coarse sources
    Sf = oc.kinterp(det["S_coarse"], oc.Cheb(1e-2, 1e4, 130, f=lambda k: np.tanh(k / 2000.0),
                                             finv=lambda y: 2000.0 * np.arctanh(y)), det["ks_fine"])
    Th = oc.los(Sf, det["taus"], det["ks_fine"], ls, oc.bessel_table(ls), [oc.INT_MAX, oc.INT_MAX, 10])
    L = ls.astype(np.float64)
    Th[0] /= det["ks_fine"][:, None]
    Th[1] *= np.sqrt((L + 2) * (L + 1) * L * (L - 1))[None, :] / det["ks_fine"][:, None] ** 2
    for s in range(3):
        scale = np.abs(Th[s]).max()
        assert np.max(np.abs(det["Th"][s] - Th[s])) <= 1e-10 * scale, s


@pytest.mark.gpu
def test_gpu_cmb_batch_of_cosmologies_matches_single_runs(fixture5):
    # several cosmologies in one source launch (each on its own τ grid) give the single-set results
    from symboltz.bessel import SphericalBesselCache
    from symboltz.cmb import spectrum_cmb, spectrum_cmb_sets
    from symboltz.problem import remake, solvebg_sets
    print("case: C_l batch of 3 cosmologies (lmax 5), 152 l", flush=True)
    S = fixture5
    d, lsc, ls = fixture_ls()
    base = CosmologyProblem(S.M, S.pars)
    probs = [base, remake(base, {"h": 0.70, "Omega_c0": 0.25}), remake(base, {"ns": 0.94, "Omega_b0": 0.047})]
    jl = SphericalBesselCache(ls)
    modes = ["TT", "EE", "TE", "ψψ"]
    bgs = solvebg_sets(probs, device=True)         # batched on the GPU
    Db = spectrum_cmb_sets(modes, probs, jl, normalization="Dl", bgs=bgs)
    for i, p in enumerate(probs):
        Ds = spectrum_cmb(modes, p, jl, normalization="Dl", bg=bgs[i])
        assert np.max(np.abs(Db[i] - Ds)) <= 1e-12 * np.abs(Ds).max(), i
    # the host background gives its own knots, hence a τ grid shifted by rounding-level step
    # differences: the spectra then differ by the grid's trapezoid error only
    Dh = spectrum_cmb(modes, base, jl, normalization="Dl")
    for m in range(4):
        err = normwise(Db[0][:, m], Dh[:, m])
        print(f"mode {modes[m]}: GPU vs host background {err:.1e}")
        assert err < 1e-3
    # a varied cosmology against the oracle pipeline on the same background knots
    from oracle import oracle
    bg = bgs[1]
    O = oracle.Oracle(S.M, bg.P, bg.t, bg.u, bg.du)
    Do, _ = oc.spectrum_cmb(O, lambda k: spectrum_primordial(k, probs[1]), modes, ls)
    for i in range(4):
        assert normwise(Db[1][:, i], Do[:, i]) < 1e-4, i


@pytest.mark.gpu
def test_gpu_source_grid_adaptive_matches_oracle(fixture5):
    from symboltz.cmb import source_grid_adaptive
    from symboltz.problem import solvebg
    from oracle import oracle
    print("case: adaptive k refinement of S_T, S_E on 30 τ, k in [1, 2000] H0/c", flush=True)
    S = fixture5
    prob = CosmologyProblem(S.M, S.pars)
    bg = solvebg(prob)
    taus = np.linspace(0.05, bg.tau0, 30)
    ks0 = np.geomspace(1.0, 2000.0, 8)
    O = oracle.Oracle(S.M, bg.P, bg.t, bg.u, bg.du)

    def solve1(k):
        r = O.solve(np.array([k]), np.array([bg.t[0]]), saveat=taus, sources=True, nthreads=1)
        return r["src"][0][:, :2]

    # absolute tolerance: 1 % of the largest source vector on the initial grid (a relative test
    # per τ would chase the tiny pre-recombination values)
    atol = 0.01 * max(np.linalg.norm(solve1(k)[:-1], axis=1).max() for k in ks0)
    tm = {}
    K, Sg = source_grid_adaptive(prob, ["ST", "SE"], taus, ks0, bg=bg, ktransform=(np.log, np.exp), atol=atol,
                                 timing=tm)
    print(f"{len(K)} wavenumbers from {tm['launches']} launches", flush=True)
    Ko, So = oc.refine_queue(solve1, ks0, np.log, np.exp, 0.0, atol, slice(0, len(taus) - 1))
    # the same decisions up to the rare interval whose test sits at the threshold within the
    # two solvers' agreement
    common = np.intersect1d(K, Ko)
    assert len(common) >= 0.97 * max(len(K), len(Ko)), (len(K), len(Ko), len(common))
    ig = np.searchsorted(K, common)
    io = np.searchsorted(Ko, common)
    scale = np.abs(So).max(axis=(0, 1))
    assert np.max(np.abs(Sg[:, ig, :] - So[io].transpose(1, 0, 2)) / scale) < 1e-4


@pytest.mark.gpu
def test_gpu_cmb_temperature_polarization_only_matches_oracle(fixture5):
    # without lensing the reference samples 61 linear Chebyshev nodes on [1e-2, 2e3]
    # (angular.jl:270-272); TT, TE, EE at the cached l only
    from symboltz.bessel import SphericalBesselCache
    from symboltz.cmb import spectrum_cmb
    print("case: C_l TT/TE/EE without lensing, 61 source modes", flush=True)
    S = fixture5
    ls = np.unique(np.round(np.geomspace(2, 1500, 60)).astype(int))
    prob = CosmologyProblem(S.M, S.pars)
    from symboltz.problem import solvebg
    bg = solvebg(prob)
    jl = SphericalBesselCache(ls)
    det = {}
    D = spectrum_cmb(["TT", "TE", "EE"], prob, jl, normalization="Dl", bg=bg, details=det)
    assert len(det["ks_coarse"]) == 61 and det["ks_coarse"][0] == 2e3
    from oracle import oracle
    O = oracle.Oracle(S.M, bg.P, bg.t, bg.u, bg.du)
    Do, _ = oc.spectrum_cmb(O, lambda k: spectrum_primordial(k, prob), ["TT", "TE", "EE"], ls, lensing=False)
    for i in range(3):
        assert normwise(D[:, i], Do[:, i]) < 1e-4, i


@pytest.mark.gpu
def test_gpu_cmb_spectra_finite_to_l3000_and_units(fixture5):
    # test/runtests.jl:561-565: TT and EE finite for l = 20:20:3000; and the μK² unit (angular.jl:318-324)
    from symboltz.bessel import SphericalBesselCache
    from symboltz.cmb import spectrum_cmb
    from symboltz.problem import solvebg
    print("case: C_l TT, EE at l = 20:20:3000; units", flush=True)
    S = fixture5
    prob = CosmologyProblem(S.M, S.pars)
    bg = solvebg(prob)
    jl = SphericalBesselCache(np.arange(20, 3001, 20))
    D = spectrum_cmb(["TT", "EE"], prob, jl, normalization="Dl", bg=bg)
    assert D.shape == (150, 2) and np.all(np.isfinite(D)) and np.all(D[:, 0] > 0)
    Du = spectrum_cmb(["TT"], prob, jl, normalization="Dl", bg=bg, unit="uK")
    assert np.allclose(Du[:, 0], D[:, 0] * (S.pars["T0"] * 1e6) ** 2, rtol=1e-12, atol=0)
    # the first acoustic peak of D_l^TT sits near l = 220 with a few thousand μK²
    lpk = jl.l[np.argmax(Du[:15, 0])]
    assert 180 <= lpk <= 260 and 3000 < Du[:, 0].max() < 8000
    with pytest.raises(ValueError):
        spectrum_cmb(["TT"], prob, jl, unit="m")
