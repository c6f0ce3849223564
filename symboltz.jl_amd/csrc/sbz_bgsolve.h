// Background + thermodynamics solve, shared by the host entry point (sbz_background.cpp) and the
// batched GPU kernel (sbz_bgsolve.hip, one thread per parameter set).
//
// Mirrors the reference's `solvebg` (src/solve.jl:427-435) with the a = 1 termination callback
// (:158-203): integrate the five background unknowns (a, X_H+, X_He+, ΔT, _κ) with Rodas5P from
// τ_bg0, stop when a = 1, record τ0, κ0 = _κ(τ0) and τrec (argmax of the visibility function over
// the stored steps), and return the step times with u and u' as the knots of the cubic Hermite
// background spline (src/utils.jl:118-127).  Everything is fixed-size: no allocation, no library
// calls beyond math functions, so the same code runs on the host and on the device.
#pragma once
#include "sbz_rodas5p.h"
#include "sbz_thermo.h"

#define SBZ_BG_NB 5

enum {
    SBZ_BG_OK = 0,
    SBZ_BG_MAXITERS = 1,     // more than 100000 step attempts
    SBZ_BG_DTMIN = 2,        // step size below 1e-14 |τ|
    SBZ_BG_MAXKNOTS = 3,     // more accepted steps than the knot buffer holds
    SBZ_BG_NOEND = 4,        // a = 1 not reached before τ = 100
};

inline const char* sbz_bg_status_text(int st) {
    switch (st) {
        case SBZ_BG_MAXITERS: return "maximum number of iterations reached";
        case SBZ_BG_DTMIN: return "step size below minimum";
        case SBZ_BG_MAXKNOTS: return "more steps than max_knots";
        case SBZ_BG_NOEND: return "a = 1 was not reached";
        default: return "success";
    }
}

struct SbzBgSys {
    const double* P;
    SBZ_HD void f(const double* u, double* du) const { sbz_bg_rhs<double>(P, u, du); }
    // exact Jacobian by forward-mode dual numbers, one column per pass
    SBZ_HD void jac(const double* u, double J[SBZ_BG_NB][SBZ_BG_NB]) const {
        Dual ud[SBZ_BG_NB], dd[SBZ_BG_NB];
        for (int j = 0; j < SBZ_BG_NB; ++j) {
            for (int i = 0; i < SBZ_BG_NB; ++i) ud[i] = Dual(u[i], i == j ? 1.0 : 0.0);
            sbz_bg_rhs<Dual>(P, ud, dd);
            for (int i = 0; i < SBZ_BG_NB; ++i) J[i][j] = dd[i].d;
        }
    }
};

// dense LU with partial pivoting of the 5 x 5 W
SBZ_HD bool sbz_bg_lu(double A[SBZ_BG_NB][SBZ_BG_NB], int piv[SBZ_BG_NB]) {
    for (int k = 0; k < SBZ_BG_NB; ++k) {
        int p = k;
        double m = fabs(A[k][k]);
        for (int i = k + 1; i < SBZ_BG_NB; ++i)
            if (fabs(A[i][k]) > m) {
                m = fabs(A[i][k]);
                p = i;
            }
        if (!(m > 0.0) || !__builtin_isfinite(m)) return false;
        piv[k] = p;
        // swap only the active columns: the stored multipliers of earlier columns stay in the
        // row order they were computed in, which is what the interleaved forward solve expects
        if (p != k)
            for (int j = k; j < SBZ_BG_NB; ++j) {
                const double t = A[k][j];
                A[k][j] = A[p][j];
                A[p][j] = t;
            }
        const double r = 1.0 / A[k][k];
        for (int i = k + 1; i < SBZ_BG_NB; ++i) {
            const double l = A[i][k] * r;
            A[i][k] = l;
            for (int j = k + 1; j < SBZ_BG_NB; ++j) A[i][j] -= l * A[k][j];
        }
    }
    return true;
}

SBZ_HD void sbz_bg_lu_solve(const double A[SBZ_BG_NB][SBZ_BG_NB], const int piv[SBZ_BG_NB], double* b) {
    for (int k = 0; k < SBZ_BG_NB; ++k) {
        if (piv[k] != k) {
            const double t = b[k];
            b[k] = b[piv[k]];
            b[piv[k]] = t;
        }
        for (int i = k + 1; i < SBZ_BG_NB; ++i) b[i] -= A[i][k] * b[k];
    }
    for (int k = SBZ_BG_NB - 1; k >= 0; --k) {
        double s = b[k];
        for (int j = k + 1; j < SBZ_BG_NB; ++j) s -= A[k][j] * b[j];
        b[k] = s / A[k][k];
    }
}

SBZ_HD double sbz_bg_err_norm(const double* e, const double* u0, const double* u1, double atol, double rtol) {
    double s = 0.0;
    for (int i = 0; i < SBZ_BG_NB; ++i) {
        const double sc = atol + fmax(fabs(u0[i]), fabs(u1[i])) * rtol;
        const double r = e[i] / sc;
        s += r * r;
    }
    return sqrt(s / SBZ_BG_NB);
}

// One Rodas5P step of size h from u; writes u1 and the error estimate.  The background ODE is
// autonomous, so the ∂f/∂t terms vanish.
SBZ_HD bool sbz_bg_step(const SbzBgSys& sys, const double* u, double h, double* u1, double* err) {
    const double A[9][8] = {
        {0}, {0},
        {SBZ_R5_A21},
        {SBZ_R5_A31, SBZ_R5_A32},
        {SBZ_R5_A41, SBZ_R5_A42, SBZ_R5_A43},
        {SBZ_R5_A51, SBZ_R5_A52, SBZ_R5_A53, SBZ_R5_A54},
        {SBZ_R5_A61, SBZ_R5_A62, SBZ_R5_A63, SBZ_R5_A64, SBZ_R5_A65},
        {SBZ_R5_A61, SBZ_R5_A62, SBZ_R5_A63, SBZ_R5_A64, SBZ_R5_A65, 1.0},
        {SBZ_R5_A61, SBZ_R5_A62, SBZ_R5_A63, SBZ_R5_A64, SBZ_R5_A65, 1.0, 1.0},
    };
    const double C[9][8] = {
        {0}, {0},
        {SBZ_R5_C21},
        {SBZ_R5_C31, SBZ_R5_C32},
        {SBZ_R5_C41, SBZ_R5_C42, SBZ_R5_C43},
        {SBZ_R5_C51, SBZ_R5_C52, SBZ_R5_C53, SBZ_R5_C54},
        {SBZ_R5_C61, SBZ_R5_C62, SBZ_R5_C63, SBZ_R5_C64, SBZ_R5_C65},
        {SBZ_R5_C71, SBZ_R5_C72, SBZ_R5_C73, SBZ_R5_C74, SBZ_R5_C75, SBZ_R5_C76},
        {SBZ_R5_C81, SBZ_R5_C82, SBZ_R5_C83, SBZ_R5_C84, SBZ_R5_C85, SBZ_R5_C86, SBZ_R5_C87},
    };
    double J[SBZ_BG_NB][SBZ_BG_NB], W[SBZ_BG_NB][SBZ_BG_NB];
    int piv[SBZ_BG_NB];
    sys.jac(u, J);
    const double ig = 1.0 / (SBZ_R5_GAMMA * h);
    for (int i = 0; i < SBZ_BG_NB; ++i)
        for (int j = 0; j < SBZ_BG_NB; ++j) W[i][j] = (i == j ? ig : 0.0) - J[i][j];
    if (!sbz_bg_lu(W, piv)) return false;
    double k[9][SBZ_BG_NB], us[SBZ_BG_NB], rhs[SBZ_BG_NB];
#pragma unroll
    for (int s = 1; s <= 8; ++s) {
        for (int i = 0; i < SBZ_BG_NB; ++i) {
            double acc = u[i];
#pragma unroll
            for (int j = 1; j < s; ++j) acc += A[s][j - 1] * k[j][i];
            us[i] = acc;
        }
        sys.f(us, rhs);
        for (int i = 0; i < SBZ_BG_NB; ++i) {
            double acc = rhs[i];
#pragma unroll
            for (int j = 1; j < s; ++j) acc += (C[s][j - 1] / h) * k[j][i];
            rhs[i] = acc;
        }
        sbz_bg_lu_solve(W, piv, rhs);
        for (int i = 0; i < SBZ_BG_NB; ++i) k[s][i] = rhs[i];
        if (s == 8)
            for (int i = 0; i < SBZ_BG_NB; ++i) {
                u1[i] = us[i] + k[8][i];
                err[i] = k[8][i];
            }
    }
    for (int i = 0; i < SBZ_BG_NB; ++i)
        if (!__builtin_isfinite(u1[i])) return false;
    return true;
}

SBZ_HD double sbz_bg_rms(const double* x, const double* sk) {
    double s = 0.0;
    for (int i = 0; i < SBZ_BG_NB; ++i) {
        const double r = x[i] / sk[i];
        s += r * r;
    }
    return sqrt(s / SBZ_BG_NB);
}

// Hairer–Wanner initial step size (as OrdinaryDiffEqCore's ode_determine_initdt)
SBZ_HD double sbz_bg_initial_dt(const SbzBgSys& sys, const double* u0, double tspan, double atol, double rtol,
                                int order) {
    double sk[SBZ_BG_NB], f0[SBZ_BG_NB], u1[SBZ_BG_NB], f1[SBZ_BG_NB], df[SBZ_BG_NB];
    for (int i = 0; i < SBZ_BG_NB; ++i) sk[i] = atol + fabs(u0[i]) * rtol;
    sys.f(u0, f0);
    const double d0 = sbz_bg_rms(u0, sk), d1 = sbz_bg_rms(f0, sk);
    double dt0 = (d0 < 1e-5 || d1 < 1e-5) ? 1e-6 : 0.01 * d0 / d1;
    dt0 = fmin(dt0, tspan);
    for (int i = 0; i < SBZ_BG_NB; ++i) u1[i] = u0[i] + dt0 * f0[i];
    sys.f(u1, f1);
    for (int i = 0; i < SBZ_BG_NB; ++i) df[i] = f1[i] - f0[i];
    const double d2 = sbz_bg_rms(df, sk) / dt0;
    const double m = fmax(d1, d2);
    const double dt1 = (m <= 1e-15) ? fmax(1e-6, dt0 * 1e-3) : pow(0.01 / m, 1.0 / (order + 1));
    return fmin(fmin(100.0 * dt0, dt1), tspan);
}

// Initial scale factor from ℋ(τ_ini) = 1/τ_ini, i.e. D(a) = a/τ (cosmologies.jl:76), solved by
// Newton's method from the radiation-era guess a = sqrt(Ω_r) τ.
SBZ_HD double sbz_bg_initial_a(const double* P, double tau) {
    // radiation-like density today: photons, massless ν and the relativistic limit of massive ν
    double Omr = P[SBZ_P_OMG] + P[SBZ_P_OMNU];
    if (P[SBZ_P_NH] > 0.0) {
        double I0 = 0.0;
        for (int i = 0; i < (int)P[SBZ_P_NX]; ++i) I0 += P[SBZ_P_W + i] * P[SBZ_P_X + i];
        Omr += P[SBZ_P_OMH] * I0 / P[SBZ_P_IRHO0];
    }
    double a = sqrt(Omr) * tau;
    for (int it = 0; it < 100; ++it) {
        Dual ad(a, 1.0);
        SbzDens<Dual> d;
        sbz_densities<Dual>(P, ad, d);
        Dual g = 1.0 / (d.Hc * tau) - 1.0;   // 1/(ℋ τ) - 1, nearly linear in a at early times
        const double step = g.v / g.d;
        a -= step;
        if (fabs(step) <= 1e-15 * fabs(a)) break;
    }
    return a;
}

// The solve.  P (in/out, SBZ_NPAR entries): derived slots are filled, then τ0, κ0 and τrec.
// Knots go to t_out [max_knots], u_out and du_out [max_knots][5]; *n_knots, *n_rejected and, on a
// failure, *t_fail (the time reached) are returned with one of the SBZ_BG_* codes.
SBZ_HD int sbz_bg_solve(double* P, double rtol, double atol, int max_knots, double* t_out, double* u_out,
                        double* du_out, int* n_knots, int* n_rejected, double* t_fail) {
    const int NB = SBZ_BG_NB;
    sbz_params_derive(P);
    SbzBgSys sys{P};
    const int order = 5;
    const int maxiters = 100000;
    double t = P[SBZ_P_TAUBG0];
    const double tmax = 100.0;
    double u[SBZ_BG_NB];
    u[0] = sbz_bg_initial_a(P, t);
    u[1] = 1.0;   // X_H+ = 1    baryons.jl:29
    u[2] = 1.0;   // X_He+ = 1   baryons.jl:33
    u[3] = 0.0;   // ΔT = 0      baryons.jl:164
    u[4] = 0.0;   // _κ = 0      baryons.jl:157
    int nk = 0, nrej = 0;
    *n_knots = 0;
    *n_rejected = 0;
    *t_fail = t;
    auto store = [&](double tt, const double* uu) -> bool {
        if (nk >= max_knots) return false;
        t_out[nk] = tt;
        double du[SBZ_BG_NB];
        sys.f(uu, du);
        for (int i = 0; i < NB; ++i) {
            u_out[nk * NB + i] = uu[i];
            du_out[nk * NB + i] = du[i];
        }
        ++nk;
        return true;
    };
    store(t, u);
    double dt = sbz_bg_initial_dt(sys, u, tmax - t, atol, rtol, order);
    double qold = SBZ_CTL_QOLDINIT;
    int iters = 0;
    bool done = false;
    int status = SBZ_BG_OK;
    while (!done) {
        if (++iters > maxiters) {
            status = SBZ_BG_MAXITERS;
            break;
        }
        if (t + dt > tmax) dt = tmax - t;
        double u1[SBZ_BG_NB], e[SBZ_BG_NB];
        const bool ok = sbz_bg_step(sys, u, dt, u1, e);
        const double EEst = ok ? sbz_bg_err_norm(e, u, u1, atol, rtol) : INFINITY;
        if (!(EEst <= 1.0)) {
            const double q11 = __builtin_isfinite(EEst) ? pow(EEst, SBZ_CTL_BETA1) : 1.0 / SBZ_CTL_QMIN;
            dt /= fmin(1.0 / SBZ_CTL_QMIN, q11 / SBZ_CTL_GAMMA);
            ++nrej;
            if (!(dt >= 1e-14 * fabs(t))) {   // also ends a NaN step size
                status = SBZ_BG_DTMIN;
                break;
            }
            continue;
        }
        double q;
        if (EEst == 0.0) {
            q = 1.0 / SBZ_CTL_QMAX;
        } else {
            const double q11 = pow(EEst, SBZ_CTL_BETA1);
            q = q11 / pow(qold, SBZ_CTL_BETA2);
            q = fmax(1.0 / SBZ_CTL_QMAX, fmin(1.0 / SBZ_CTL_QMIN, q / SBZ_CTL_GAMMA));
        }
        qold = fmax(EEst, SBZ_CTL_QOLDINIT);
        const double dtnew = dt / q;
        if (u1[0] >= 1.0) {
            // Event a = 1 inside (t, t+dt]: locate τ0 on the cubic Hermite interpolant of a
            // between the step ends, then step exactly to τ0.
            double f0[SBZ_BG_NB], f1[SBZ_BG_NB];
            sys.f(u, f0);
            sys.f(u1, f1);
            double lo = 0.0, hi = 1.0;
            for (int it = 0; it < 200; ++it) {
                const double th = 0.5 * (lo + hi);
                const double h00 = (1 + 2 * th) * (1 - th) * (1 - th), h10 = th * (1 - th) * (1 - th);
                const double h01 = th * th * (3 - 2 * th), h11 = th * th * (th - 1);
                const double av = h00 * u[0] + h10 * dt * f0[0] + h01 * u1[0] + h11 * dt * f1[0];
                if (av >= 1.0) hi = th;
                else lo = th;
                if (hi - lo < 1e-16) break;
            }
            double hev = hi * dt;
            double ue[SBZ_BG_NB], ee[SBZ_BG_NB];
            if (hev > 0.0 && sbz_bg_step(sys, u, hev, ue, ee)) {
                // one Newton refinement on the actual step: a(τ) - 1 with da/dτ = a ℋ
                double fe[SBZ_BG_NB];
                sys.f(ue, fe);
                const double corr = (ue[0] - 1.0) / fe[0];
                if (fabs(corr) < hev) {
                    hev -= corr;
                    sbz_bg_step(sys, u, hev, ue, ee);
                }
            } else {
                hev = dt;
                for (int i = 0; i < NB; ++i) ue[i] = u1[i];
            }
            t = t + hev;
            for (int i = 0; i < NB; ++i) u[i] = ue[i];
            if (!store(t, u)) {
                status = SBZ_BG_MAXKNOTS;
                break;
            }
            done = true;
            break;
        }
        t += dt;
        for (int i = 0; i < NB; ++i) u[i] = u1[i];
        if (!store(t, u)) {
            status = SBZ_BG_MAXKNOTS;
            break;
        }
        if (t >= tmax) break;
        dt = dtnew;
    }
    *n_knots = nk;
    *n_rejected = nrej;
    *t_fail = t;
    if (status != SBZ_BG_OK) return status;
    if (!done) return SBZ_BG_NOEND;
    // Knot derivatives.  The reference takes u' from the integrator's continuous extension
    // (sol(ts, Val{1}), src/utils.jl:123), which is smooth.  For the stiff thermodynamic unknowns
    // X_H+, X_He+ and ΔT, f(u) is not: they sit on a slow manifold where J ~ 1e20 amplifies the
    // solver's tolerance-level deviations into derivatives that are wrong by orders of magnitude.
    // Their u' is therefore the derivative of the Lagrange polynomial through the five nearest
    // knots, like a continuous extension; a and _κ keep the exact u' = f(u).  The stencil reads
    // only u, so du is overwritten in place.
    for (int i = 0; i < nk; ++i) {
        int lo = i - 2 < nk - 5 ? i - 2 : nk - 5;
        lo = lo > 0 ? lo : 0;
        const int hi = nk - 1 < lo + 4 ? nk - 1 : lo + 4;
        double w[5] = {0, 0, 0, 0, 0};
        const double ti = t_out[i];
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
            du_out[i * NB + c] = s;
        }
    }
    // τ0, κ0 and τrec (solve.jl:181-195)
    P[SBZ_P_TAU0] = t_out[nk - 1];
    P[SBZ_P_KAPPA0] = u_out[(nk - 1) * NB + 4];
    double vmax = -INFINITY, trec = t_out[0];
    for (int i = 0; i < nk; ++i) {
        const double kd = du_out[i * NB + 4];
        const double v = -kd * exp(-(u_out[i * NB + 4] - P[SBZ_P_KAPPA0]));
        if (v > vmax) {
            vmax = v;
            trec = t_out[i];
        }
    }
    P[SBZ_P_TAUREC] = trec;
    return SBZ_BG_OK;
}
