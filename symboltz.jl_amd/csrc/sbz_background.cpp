// Host-side background + thermodynamics solve (the input stage of the perturbation hot path).
//
// Mirrors the reference's `solvebg` (src/solve.jl:427-435) with the a = 1 termination callback
// (:158-203): integrate the five background unknowns (a, X_H+, X_He+, ΔT, _κ) with Rodas5P at
// rtol = atol = 1e-7 from τ_bg0 = 1e-6, stop when a = 1, record τ0, κ0 = _κ(τ0) and τrec
// (argmax of the visibility function over the stored steps), and return the step times with
// u and u' = f(τ, u) as the knots of the cubic Hermite background spline (src/utils.jl:118-127).
#include <hip/hip_runtime.h>
#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#include "sbz_rodas5p.h"
#include "sbz_thermo.h"
#include "sbz_internal.h"

namespace {

constexpr int NB = 5;

struct BgSys {
    const double* P;
    void f(const double* u, double* du) const { sbz_bg_rhs<double>(P, u, du); }
    // Exact Jacobian by forward-mode dual numbers, one column per pass.
    void jac(const double* u, double J[NB][NB]) const {
        Dual ud[NB], dd[NB];
        for (int j = 0; j < NB; ++j) {
            for (int i = 0; i < NB; ++i) ud[i] = Dual(u[i], i == j ? 1.0 : 0.0);
            sbz_bg_rhs<Dual>(P, ud, dd);
            for (int i = 0; i < NB; ++i) J[i][j] = dd[i].d;
        }
    }
};

// Dense LU with partial pivoting for the 5x5 W matrix.
bool lu_factor(double A[NB][NB], int piv[NB]) {
    for (int k = 0; k < NB; ++k) {
        int p = k;
        double m = std::fabs(A[k][k]);
        for (int i = k + 1; i < NB; ++i)
            if (std::fabs(A[i][k]) > m) { m = std::fabs(A[i][k]); p = i; }
        if (!(m > 0.0) || !std::isfinite(m)) return false;
        piv[k] = p;
        // swap only the active columns: the stored multipliers of earlier columns stay in the
        // row order they were computed in, which is what the interleaved forward solve expects
        if (p != k)
            for (int j = k; j < NB; ++j) std::swap(A[k][j], A[p][j]);
        double r = 1.0 / A[k][k];
        for (int i = k + 1; i < NB; ++i) {
            double l = A[i][k] * r;
            A[i][k] = l;
            for (int j = k + 1; j < NB; ++j) A[i][j] -= l * A[k][j];
        }
    }
    return true;
}

void lu_solve(const double A[NB][NB], const int piv[NB], double* b) {
    for (int k = 0; k < NB; ++k) {
        if (piv[k] != k) std::swap(b[k], b[piv[k]]);
        for (int i = k + 1; i < NB; ++i) b[i] -= A[i][k] * b[k];
    }
    for (int k = NB - 1; k >= 0; --k) {
        double s = b[k];
        for (int j = k + 1; j < NB; ++j) s -= A[k][j] * b[j];
        b[k] = s / A[k][k];
    }
}

double rms_residual(const double* e, const double* u0, const double* u1, double atol, double rtol) {
    double s = 0.0;
    for (int i = 0; i < NB; ++i) {
        double sc = atol + std::fmax(std::fabs(u0[i]), std::fabs(u1[i])) * rtol;
        double r = e[i] / sc;
        s += r * r;
    }
    return std::sqrt(s / NB);
}

// One Rodas5P step of size h from (t, u); writes u1 and the error estimate.  The background
// ODE is autonomous, so the ∂f/∂t terms vanish.
bool rodas5p_step(const BgSys& sys, double t, const double* u, double h, double* u1, double* err) {
    (void)t;
    static const double A[9][8] = {
        {0}, {0},
        {SBZ_R5_A21},
        {SBZ_R5_A31, SBZ_R5_A32},
        {SBZ_R5_A41, SBZ_R5_A42, SBZ_R5_A43},
        {SBZ_R5_A51, SBZ_R5_A52, SBZ_R5_A53, SBZ_R5_A54},
        {SBZ_R5_A61, SBZ_R5_A62, SBZ_R5_A63, SBZ_R5_A64, SBZ_R5_A65},
        {SBZ_R5_A61, SBZ_R5_A62, SBZ_R5_A63, SBZ_R5_A64, SBZ_R5_A65, 1.0},
        {SBZ_R5_A61, SBZ_R5_A62, SBZ_R5_A63, SBZ_R5_A64, SBZ_R5_A65, 1.0, 1.0},
    };
    static const double C[9][8] = {
        {0}, {0},
        {SBZ_R5_C21},
        {SBZ_R5_C31, SBZ_R5_C32},
        {SBZ_R5_C41, SBZ_R5_C42, SBZ_R5_C43},
        {SBZ_R5_C51, SBZ_R5_C52, SBZ_R5_C53, SBZ_R5_C54},
        {SBZ_R5_C61, SBZ_R5_C62, SBZ_R5_C63, SBZ_R5_C64, SBZ_R5_C65},
        {SBZ_R5_C71, SBZ_R5_C72, SBZ_R5_C73, SBZ_R5_C74, SBZ_R5_C75, SBZ_R5_C76},
        {SBZ_R5_C81, SBZ_R5_C82, SBZ_R5_C83, SBZ_R5_C84, SBZ_R5_C85, SBZ_R5_C86, SBZ_R5_C87},
    };
    double J[NB][NB], W[NB][NB];
    int piv[NB];
    sys.jac(u, J);
    double ig = 1.0 / (SBZ_R5_GAMMA * h);
    for (int i = 0; i < NB; ++i)
        for (int j = 0; j < NB; ++j) W[i][j] = (i == j ? ig : 0.0) - J[i][j];
    if (!lu_factor(W, piv)) return false;
    double k[9][NB], us[NB], rhs[NB];
    for (int s = 1; s <= 8; ++s) {
        for (int i = 0; i < NB; ++i) {
            double acc = u[i];
            for (int j = 1; j < s; ++j) acc += A[s][j - 1] * k[j][i];
            us[i] = acc;
        }
        sys.f(us, rhs);
        for (int i = 0; i < NB; ++i) {
            double acc = rhs[i];
            for (int j = 1; j < s; ++j) acc += (C[s][j - 1] / h) * k[j][i];
            rhs[i] = acc;
        }
        lu_solve(W, piv, rhs);
        for (int i = 0; i < NB; ++i) k[s][i] = rhs[i];
        if (s == 8) {
            for (int i = 0; i < NB; ++i) {
                u1[i] = us[i] + k[8][i];
                err[i] = k[8][i];
            }
        }
    }
    for (int i = 0; i < NB; ++i)
        if (!std::isfinite(u1[i])) return false;
    return true;
}

double rms(const double* x, const double* sk) {
    double s = 0.0;
    for (int i = 0; i < NB; ++i) { double r = x[i] / sk[i]; s += r * r; }
    return std::sqrt(s / NB);
}

// Hairer–Wanner initial step size (as OrdinaryDiffEqCore's ode_determine_initdt).
double initial_dt(const BgSys& sys, double t0, const double* u0, double tspan, double atol, double rtol, int order) {
    double sk[NB], f0[NB], u1[NB], f1[NB], df[NB];
    for (int i = 0; i < NB; ++i) sk[i] = atol + std::fabs(u0[i]) * rtol;
    sys.f(u0, f0);
    double d0 = rms(u0, sk), d1 = rms(f0, sk);
    double dt0 = (d0 < 1e-5 || d1 < 1e-5) ? 1e-6 : 0.01 * d0 / d1;
    dt0 = std::fmin(dt0, tspan);
    for (int i = 0; i < NB; ++i) u1[i] = u0[i] + dt0 * f0[i];
    sys.f(u1, f1);
    for (int i = 0; i < NB; ++i) df[i] = f1[i] - f0[i];
    double d2 = rms(df, sk) / dt0;
    double m = std::fmax(d1, d2);
    double dt1 = (m <= 1e-15) ? std::fmax(1e-6, dt0 * 1e-3) : std::pow(0.01 / m, 1.0 / (order + 1));
    (void)t0;
    return std::fmin(std::fmin(100.0 * dt0, dt1), tspan);
}

}  // namespace

// Initial scale factor from ℋ(τ_ini) = 1/τ_ini, i.e. D(a) = a/τ (cosmologies.jl:76), solved by
// Newton's method from the radiation-era guess a = sqrt(Ω_r) τ.
static double initial_scale_factor(const double* P, double tau) {
    // radiation-like density today: photons, massless ν and the relativistic limit of massive ν
    double Omr = P[SBZ_P_OMG] + P[SBZ_P_OMNU];
    if (P[SBZ_P_NH] > 0.0) {
        double I0 = 0.0;
        for (int i = 0; i < (int)P[SBZ_P_NX]; ++i) I0 += P[SBZ_P_W + i] * P[SBZ_P_X + i];
        Omr += P[SBZ_P_OMH] * I0 / P[SBZ_P_IRHO0];
    }
    double a = std::sqrt(Omr) * tau;
    for (int it = 0; it < 100; ++it) {
        Dual ad(a, 1.0);
        SbzDens<Dual> d;
        sbz_densities<Dual>(P, ad, d);
        Dual g = 1.0 / (d.Hc * tau) - 1.0;  // 1/(ℋ τ) - 1, nearly linear in a at early times
        double step = g.v / g.d;
        a -= step;
        if (std::fabs(step) <= 1e-15 * std::fabs(a)) break;
    }
    return a;
}

// Background right-hand side at one state (host helper for tests and the Python mirror).
extern "C" int sbz_background_rhs(const double* params, const double* u, double* du) {
    if (!params || !u || !du) return SBZ_ERR_ARG;
    sbz_bg_rhs<double>(params, u, du);
    return SBZ_OK;
}

extern "C" int sbz_background_solve(double* params, double rtol, double atol, int max_knots,
                                    double* t_out, double* u_out, double* du_out, int* n_knots,
                                    int* n_rejected) {
    sbz_clear_error();
    if (!params || !t_out || !u_out || !du_out || !n_knots || max_knots < 3) {
        sbz_set_error("sbz_background_solve: null argument or max_knots < 3");
        return SBZ_ERR_ARG;
    }
    double* P = params;
    sbz_params_derive(P);
    BgSys sys{P};
    const int order = 5;
    const int maxiters = 100000;
    double t = P[SBZ_P_TAUBG0];
    const double tmax = 100.0;
    double u[NB];
    u[0] = initial_scale_factor(P, t);
    u[1] = 1.0;  // X_H+ = 1    baryons.jl:29
    u[2] = 1.0;  // X_He+ = 1   baryons.jl:33
    u[3] = 0.0;  // ΔT = 0      baryons.jl:164
    u[4] = 0.0;  // _κ = 0      baryons.jl:157
    int nk = 0;
    auto store = [&](double tt, const double* uu) -> bool {
        if (nk >= max_knots) return false;
        t_out[nk] = tt;
        double du[NB];
        sys.f(uu, du);
        for (int i = 0; i < NB; ++i) {
            u_out[nk * NB + i] = uu[i];
            du_out[nk * NB + i] = du[i];
        }
        ++nk;
        return true;
    };
    store(t, u);
    double dt = initial_dt(sys, t, u, tmax - t, atol, rtol, order);
    double qold = SBZ_CTL_QOLDINIT;
    int iters = 0, nrej = 0;
    bool done = false;
    while (!done) {
        if (++iters > maxiters) {
            sbz_set_error("background: maximum number of iterations reached");
            *n_knots = nk;
            return SBZ_ERR_SOLVE;
        }
        if (t + dt > tmax) dt = tmax - t;
        double u1[NB], e[NB];
        bool ok = rodas5p_step(sys, t, u, dt, u1, e);
        double EEst = ok ? rms_residual(e, u, u1, atol, rtol) : INFINITY;
        if (!(EEst <= 1.0)) {
            double q11 = std::isfinite(EEst) ? std::pow(EEst, SBZ_CTL_BETA1) : 1.0 / SBZ_CTL_QMIN;
            dt /= std::fmin(1.0 / SBZ_CTL_QMIN, q11 / SBZ_CTL_GAMMA);
            ++nrej;
            if (dt < 1e-14 * std::fabs(t)) {
                sbz_set_error("background: step size below minimum at tau=%g (a=%g XH=%g XHe=%g dT=%g kappa=%g, EEst=%g)",
                              t, u[0], u[1], u[2], u[3], u[4], EEst);
                *n_knots = nk;
                return SBZ_ERR_SOLVE;
            }
            continue;
        }
        double q;
        if (EEst == 0.0) {
            q = 1.0 / SBZ_CTL_QMAX;
        } else {
            double q11 = std::pow(EEst, SBZ_CTL_BETA1);
            q = q11 / std::pow(qold, SBZ_CTL_BETA2);
            q = std::fmax(1.0 / SBZ_CTL_QMAX, std::fmin(1.0 / SBZ_CTL_QMIN, q / SBZ_CTL_GAMMA));
        }
        qold = std::fmax(EEst, SBZ_CTL_QOLDINIT);
        double dtnew = dt / q;
        if (u1[0] >= 1.0) {
            // Event a = 1 inside (t, t+dt]: locate τ0 on the cubic Hermite interpolant of a
            // between the step ends, then step exactly to τ0.
            double f0[NB], f1[NB];
            sys.f(u, f0);
            sys.f(u1, f1);
            double lo = 0.0, hi = 1.0;
            for (int it = 0; it < 200; ++it) {
                double th = 0.5 * (lo + hi);
                double h00 = (1 + 2 * th) * (1 - th) * (1 - th), h10 = th * (1 - th) * (1 - th);
                double h01 = th * th * (3 - 2 * th), h11 = th * th * (th - 1);
                double av = h00 * u[0] + h10 * dt * f0[0] + h01 * u1[0] + h11 * dt * f1[0];
                if (av >= 1.0) hi = th; else lo = th;
                if (hi - lo < 1e-16) break;
            }
            double hev = hi * dt;
            double ue[NB], ee[NB];
            if (hev > 0.0 && rodas5p_step(sys, t, u, hev, ue, ee)) {
                // one Newton refinement on the actual step: a(τ) - 1 with da/dτ = a ℋ
                double fe[NB];
                sys.f(ue, fe);
                double corr = (ue[0] - 1.0) / fe[0];
                if (std::fabs(corr) < hev) {
                    hev -= corr;
                    rodas5p_step(sys, t, u, hev, ue, ee);
                }
            } else {
                hev = dt;
                std::memcpy(ue, u1, sizeof(ue));
            }
            t = t + hev;
            std::memcpy(u, ue, sizeof(u));
            if (!store(t, u)) break;
            done = true;
            break;
        }
        t += dt;
        std::memcpy(u, u1, sizeof(u));
        if (!store(t, u)) {
            sbz_set_error("background: more steps than max_knots");
            *n_knots = nk;
            return SBZ_ERR_SOLVE;
        }
        if (t >= tmax) break;
        dt = dtnew;
    }
    *n_knots = nk;
    if (n_rejected) *n_rejected = nrej;
    if (!done) {
        sbz_set_error("background: a = 1 was not reached");
        return SBZ_ERR_SOLVE;
    }
    // Knot derivatives.  The reference takes u' from the integrator's continuous extension
    // (sol(ts, Val{1}), src/utils.jl:123), which is smooth.  For the stiff thermodynamic unknowns
    // X_H+, X_He+ and ΔT, f(u) is not: they sit on a slow manifold where J ~ 1e20 amplifies the
    // solver's tolerance-level deviations into derivatives that are wrong by orders of magnitude.
    // Their u' is therefore taken from the derivative of the Lagrange polynomial through the five
    // nearest knots, like a continuous extension; a and _κ keep the exact u' = f(u).
    {
        std::vector<double> dtmp((size_t)nk * NB);
        for (int i = 0; i < nk; ++i) {
            int lo = std::max(0, std::min(i - 2, nk - 5));
            int hi = std::min(nk - 1, lo + 4);
            double w[5] = {0, 0, 0, 0, 0};
            double ti = t_out[i];
            for (int j = lo; j <= hi; ++j) {
                double wj;
                if (j == i) {
                    wj = 0.0;
                    for (int m = lo; m <= hi; ++m)
                        if (m != i) wj += 1.0 / (ti - t_out[m]);
                } else {
                    double num = 1.0, den = 1.0;
                    for (int m = lo; m <= hi; ++m) {
                        if (m == j) continue;
                        den *= (t_out[j] - t_out[m]);
                        if (m != i) num *= (ti - t_out[m]);
                    }
                    wj = num / den;
                }
                w[j - lo] = wj;
            }
            for (int c = 1; c <= 3; ++c) {
                double s = 0.0;
                for (int j = lo; j <= hi; ++j) s += w[j - lo] * u_out[j * NB + c];
                dtmp[(size_t)i * NB + c] = s;
            }
        }
        for (int i = 0; i < nk; ++i)
            for (int c = 1; c <= 3; ++c) du_out[i * NB + c] = dtmp[(size_t)i * NB + c];
    }
    // τ0, κ0 and τrec (solve.jl:181-195)
    P[SBZ_P_TAU0] = t_out[nk - 1];
    P[SBZ_P_KAPPA0] = u_out[(nk - 1) * NB + 4];
    double vmax = -INFINITY, trec = t_out[0];
    for (int i = 0; i < nk; ++i) {
        double kd = du_out[i * NB + 4];
        double v = -kd * std::exp(-(u_out[i * NB + 4] - P[SBZ_P_KAPPA0]));
        if (v > vmax) { vmax = v; trec = t_out[i]; }
    }
    P[SBZ_P_TAUREC] = trec;
    return SBZ_OK;
}
