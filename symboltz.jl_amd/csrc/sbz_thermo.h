// Background and thermodynamics of the ΛCDM / w0waCDM models, evaluated on host or device.
//
// Everything here is templated on the scalar type (double or Dual) so the same code gives the
// background ODE right-hand side for the host background solver, the background quantities the
// perturbation right-hand side reads at every stage, and their exact τ-derivatives (∂f/∂τ).
//
// The perturbation stage reads the five splined background unknowns (a, X_H+, X_He+, ΔT, _κ)
// and recomputes every other background quantity from them, as the reference does after its
// spline pass (src/utils.jl:184-236; paper/aa57450-25.tex:463-466).
#pragma once
#include "sbz_dual.h"
#include "../../include/sbz_params.h"

namespace sbzc {
// CODATA 2018 (src/constants.jl:3-10) and astronomical units (:11-19).
static constexpr double c_ = 299792458.0;
static constexpr double hP = 6.62607015e-34;
static constexpr double hbar = 1.0545718176461565e-34;
static constexpr double kB = 1.380649e-23;
static constexpr double GN = 6.6743e-11;
static constexpr double sigmaSB = 5.670374419184429e-8;
static constexpr double aR = 4.0 / c_ * sigmaSB;
static constexpr double sigmaT = 6.6524587321e-29;
static constexpr double Mpc = 3.0856775814913673e22;
static constexpr double H100 = 100.0 * 1000.0 / Mpc;
static constexpr double eV = 1.602176634e-19;
static constexpr double me = 9.1093837015e-31;
static constexpr double amu = 1.6605390666e-27;
static constexpr double mH = 1.008 * amu;        // PeriodicTable H atomic mass (constants.jl:22)
static constexpr double mHe = 4.002602 * amu;    // PeriodicTable He atomic mass (constants.jl:23)
static constexpr double PI = 3.141592653589793;

// Atomic transition energies (src/constants.jl:26-43): E = h c / λ.
static constexpr double EH_inf_1s = hP * c_ / 91.17534e-9;
static constexpr double EH_2s_1s = hP * c_ / 121.56700e-9;
static constexpr double EH_inf_2s = EH_inf_1s - EH_2s_1s;
static constexpr double lamH_2s_1s = 121.56700e-9;
static constexpr double EHe_inf_1s = hP * c_ / 50.42590e-9;
static constexpr double EHe_2s_1s = hP * c_ / 60.14045e-9;
static constexpr double EHe_2p_1s = hP * c_ / 58.43344e-9;
static constexpr double lamHe_2p_1s = 58.43344e-9;
static constexpr double fHe_2p_1s = c_ / 58.43344e-9;
static constexpr double EHe_2p_2s = EHe_2p_1s - EHe_2s_1s;
static constexpr double EHe_inf_2s = EHe_inf_1s - EHe_2s_1s;
static constexpr double EHep_inf_1s = 54.4178 * eV;
static constexpr double EHet_inf_2s = hP * c_ / 260.0463e-9;
static constexpr double lamHet_2p_1s = 59.1411e-9;
static constexpr double fHet_2p_1s = c_ / 59.1411e-9;
static constexpr double EHet_2p_1s = hP * c_ / 59.1411e-9;
static constexpr double EHet_2s_1s = hP * c_ / 62.5563e-9;
static constexpr double EHet_2p_2s = EHet_2p_1s - EHet_2s_1s;
}  // namespace sbzc

// smoothifelse(x, v1, v2; k) = ½((v1+v2) + (v2-v1) tanh(k x))            src/utils.jl:37
template <class T>
SBZ_HD T sbz_smoothifelse(T x, T v1, T v2, double k) {
    return 0.5 * ((v1 + v2) + (v2 - v1) * sbz_tanh(k * x));
}

// Massive-neutrino momentum integrals at scale factor a (src/models/neutrinos.jl:100-105).
template <class T>
SBZ_HD void sbz_nu_integrals(const double* P, T a, T* E, T& Irho, T& IP) {
    int nx = (int)P[SBZ_P_NX];
    T y = P[SBZ_P_Y0H] * a;
    Irho = T(0.0);
    IP = T(0.0);
    // fixed trip count with a predicate: E[] stays in registers on the device
#pragma unroll
    for (int i = 0; i < SBZ_MAX_NX; ++i) {
        if (i < nx) {
            double x = P[SBZ_P_X + i];
            T Ei = sbz_sqrt(x * x + y * y);
            E[i] = Ei;
            Irho = Irho + P[SBZ_P_W + i] * Ei;
            IP = IP + P[SBZ_P_W + i] * (x * x / Ei);
        } else {
            E[i] = T(1.0);
        }
    }
}

// Densities ρ_s(a) in the reference's units ρ = 3/(8π) Ω (generic_species.jl:23-24,
// neutrinos.jl:103-104, dark_energy.jl:47), and the conformal Hubble rate from the Friedmann
// equation D(a) = sqrt(8π/3 ρ) a² (gravity.jl:31), ℋ = D(a)/a (metric.jl:26).
template <class T>
struct SbzDens {
    T rho_g, rho_nu, rho_c, rho_b, rho_h, P_h, Irho, IP, rho_de, w_de, rho_tot, Hc;
    T E[SBZ_MAX_NX];
};

template <class T>
SBZ_HD void sbz_densities(const double* P, T a, SbzDens<T>& d) {
    const double k38 = 3.0 / (8.0 * sbzc::PI);
    T a2 = a * a, a3 = a2 * a, a4 = a2 * a2;
    d.rho_g = k38 * P[SBZ_P_OMG] / a4;
    d.rho_nu = k38 * P[SBZ_P_OMNU] / a4;
    d.rho_c = k38 * P[SBZ_P_OMC] / a3;
    d.rho_b = k38 * P[SBZ_P_OMB] / a3;
    if (P[SBZ_P_NH] > 0.0) {
        sbz_nu_integrals(P, a, d.E, d.Irho, d.IP);
        T f = k38 * P[SBZ_P_OMH] / (P[SBZ_P_IRHO0] * a4);
        d.rho_h = f * d.Irho;
        d.P_h = f * d.IP / 3.0;
    } else {
        d.rho_h = T(0.0);
        d.P_h = T(0.0);
        d.Irho = T(1.0);
        d.IP = T(0.0);
    }
    if (P[SBZ_P_DE] == 1.0) {
        double w0 = P[SBZ_P_W0], wa = P[SBZ_P_WA];
        // ρ = 3/(8π) Ω0 |a|^(-3(1+w0+wa)) exp(-3 wa (1-a))                  dark_energy.jl:47
        d.rho_de = k38 * P[SBZ_P_OMDE] * sbz_pow(sbz_abs(a), -3.0 * (1.0 + w0 + wa)) *
                   sbz_exp(-3.0 * wa * (1.0 - a));
        d.w_de = w0 + wa * (1.0 - a);
    } else {
        d.rho_de = T(k38 * P[SBZ_P_OMDE]);
        d.w_de = T(-1.0);
    }
    d.rho_tot = d.rho_g + d.rho_nu + d.rho_c + d.rho_b + d.rho_h + d.rho_de;
    d.Hc = sbz_sqrt(8.0 * sbzc::PI / 3.0 * d.rho_tot) * a;
}

// Thermodynamic quantities from the splined unknowns (baryons.jl:145-212, recfast :14-116).
template <class T>
struct SbzThermo {
    T Tg, Tb, DT, Xe, nH, ne, kappadot, cs2, XHepp;
};

template <class T>
SBZ_HD T sbz_alphaHe_fit(T Tb, double q, double p) {
    // αHefit(T; q, p, T1 = 10^5.114, T2 = 3)                                 baryons.jl:40
    const double T1 = 130016.55150666007;  // 10^5.114
    const double T2 = 3.0;
    T s2 = sbz_sqrt(Tb / T2), s1 = sbz_sqrt(Tb / T1);
    return q / (s2 * sbz_pow(1.0 + s2, 1.0 - p) * sbz_pow(1.0 + s1, 1.0 + p));
}

// Fill the derived entries of a parameter vector (SBZ_P_RE*): parameter-only powers of the
// reionization terms, evaluated once per set instead of at every background evaluation.
SBZ_HD void sbz_params_derive(double* P) {
    const double z1 = P[SBZ_P_ZRE1], n1 = P[SBZ_P_NRE1], z2 = P[SBZ_P_ZRE2], n2 = P[SBZ_P_NRE2];
    P[SBZ_P_RE1F] = pow(1.0 + z1, n1);
    P[SBZ_P_RE1K] = 1.0 / (n1 * pow(1.0 + z1, n1 - 1.0) * P[SBZ_P_DZRE1]);
    P[SBZ_P_RE2F] = pow(1.0 + z2, n2);
    P[SBZ_P_RE2K] = 1.0 / (n2 * pow(1.0 + z2, n2 - 1.0) * P[SBZ_P_DZRE2]);
}

// smoothifelse(f(z_re) - f(z), 0, Xemax; k = 1/(n (1+z_re)^(n-1) Δz)), f(z) = (1+z)^n, with
// fre = f(z_re) and kk precomputed (sbz_params_derive); opz = 1 + z = 1/a    baryons.jl:128-138
template <class T>
SBZ_HD T sbz_reion_xe(T opz, double fre, double kk, double n, double xemax) {
    T fz = (n == 1.5) ? sbz_sqrt(opz) * opz : (n == 1.0) ? opz : sbz_pow(opz, n);
    return sbz_smoothifelse(fre - fz, T(0.0), T(xemax), kk);
}

template <class T>
SBZ_HD void sbz_thermo(const double* P, T a, T XH, T XHe, T dT, const SbzDens<T>& d,
                       SbzThermo<T>& th) {
    using namespace sbzc;
    const double h = P[SBZ_P_H], YHe = P[SBZ_P_YHE], fHe = P[SBZ_P_FHE];
    th.Tg = P[SBZ_P_T0] / a;                                 // T = T0/a        generic_species.jl:85
    th.Tb = dT + th.Tg;                                      // T = ΔT + Tγ     baryons.jl:188
    const double H0SI = H100 * h;
    th.nH = (1.0 - YHe) * d.rho_b * (H0SI * H0SI / GN / mH);  // baryons.jl:190
    // He++ from the quadratic Saha equation                                  baryons.jl:64-67
    T beta = 1.0 / (kB * th.Tb);
    T lam_e = hP / sbz_sqrt(2.0 * PI * me / beta);
    T lam3 = lam_e * lam_e * lam_e;
    T R = sbz_exp(-beta * EHep_inf_1s) / (th.nH * lam3);
    T den = 1.0 + fHe + R;
    th.XHepp = 2.0 * R * fHe / den / (1.0 + sbz_sqrt(1.0 + 4.0 * R * fHe / (den * den)));
    th.Xe = XH + fHe * XHe + th.XHepp;                       // baryons.jl:69
    if (P[SBZ_P_REION] != 0.0) {
        const T opz = 1.0 / a;
        th.Xe = th.Xe + sbz_reion_xe(opz, P[SBZ_P_RE1F], P[SBZ_P_RE1K], P[SBZ_P_NRE1], 1.0 + fHe);
        th.Xe = th.Xe + sbz_reion_xe(opz, P[SBZ_P_RE2F], P[SBZ_P_RE2K], P[SBZ_P_NRE2], fHe);
    }
    th.ne = th.Xe * th.nH;                                   // baryons.jl:192
    th.kappadot = -(a / H0SI) * th.ne * (sigmaT * c_);       // D(_κ)           baryons.jl:175
    // Baryon temperature derivative and sound speed                          baryons.jl:182-188
    T Tg2 = th.Tg * th.Tg;
    th.DT = -2.0 * th.Tb * d.Hc -
            (a / h) * (8.0 / 3.0 * sigmaT * aR / H100) * (Tg2 * Tg2 / (me * c_)) *
                (th.Xe / (1.0 + fHe + th.Xe)) * dT;
    T mu_c2 = (mH * c_ * c_) / (1.0 + (mH / mHe - 1.0) * YHe + th.Xe * (1.0 - YHe));
    th.cs2 = kB / mu_c2 * (th.Tb - th.DT / (3.0 * d.Hc));
}

// Right-hand side of the background ODE for u = (a, X_H+, X_He+, ΔT, _κ).
// D(a) gravity.jl:31; D(X_H+) baryons.jl:51; D(X_He+) :60 (+ triplet :111); D(ΔT) :187;
// D(_κ) :175.
template <class T>
SBZ_HD void sbz_bg_rhs(const double* P, const T* u, T* du) {
    using namespace sbzc;
    const double h = P[SBZ_P_H], fHe = P[SBZ_P_FHE];
    T a = u[0], XH = u[1], XHe = u[2], dT = u[3];
    SbzDens<T> d;
    sbz_densities(P, a, d);
    SbzThermo<T> th;
    sbz_thermo(P, a, XH, XHe, dT, d, th);
    const double H0SI = H100 * h;
    T Tb = th.Tb;
    T beta = 1.0 / (kB * Tb);
    T lam_e = hP / sbz_sqrt(2.0 * PI * me / beta);
    T lam3 = lam_e * lam_e * lam_e;
    T HSI = H0SI * d.Hc / a;  // H = H100 h ℋ/a                               baryons.jl:46
    T nH = th.nH, nHe = fHe * th.nH, ne = th.ne;
    T pref = -a / H0SI;

    // H+ + e- recombination                                                  baryons.jl:47-51,75-81
    const double LamH = 8.2245809, LamHe = 51.3, XlimC = 0.99;
    T t4 = Tb / 1e4;
    T alphaH = P[SBZ_P_FH] * 1e-19 * 4.309 * sbz_pow(t4, -0.6166) / (1.0 + 0.6703 * sbz_pow(t4, 0.5300));
    T betaH = alphaH / lam3 * sbz_exp(-beta * EH_inf_2s);
    T la = sbz_log(a);
    T g1 = (la + 7.28) / 0.18, g2 = (la + 6.73) / 0.33;
    T KHfit = (P[SBZ_P_HSW] == 1.0) ? 1.0 + (-0.14) * sbz_exp(-(g1 * g1)) + 0.079 * sbz_exp(-(g2 * g2)) : T(1.0);
    T KH = KHfit / (8.0 * PI) * (lamH_2s_1s * lamH_2s_1s * lamH_2s_1s) / HSI;
    T omXH = 1.0 - XH;
    T CHv = (1.0 + KH * LamH * nH * omXH) / (1.0 + KH * (LamH + betaH) * nH * omXH);
    T CH = sbz_smoothifelse(XH - XlimC, CHv, T(1.0), 1e3);
    du[1] = pref * CH * (alphaH * XH * ne - betaH * omXH * sbz_exp(-beta * EH_2s_1s));

    // He+ + e- singlet recombination                                         baryons.jl:54-60
    T alphaHe = sbz_alphaHe_fit(Tb, pow(10.0, -16.744), 0.711);
    T betaHe = 4.0 * alphaHe / lam3 * sbz_exp(-beta * EHe_inf_2s);
    T invK0 = 8.0 * PI * HSI / (lamHe_2p_1s * lamHe_2p_1s * lamHe_2p_1s);
    T omXHe = 1.0 - XHe;
    T invK1 = T(0.0), invK2 = T(0.0), DXHet = T(0.0);
    if (P[SBZ_P_HESW] == 6.0) {
        // Helium corrections (Heswitch = 6)                                  baryons.jl:93-113
        const double eps = 1e-9, A2ps = 1.798287e9, A2pt = 177.58;
        T tauHe = 3.0 * A2ps * nHe * (omXHe + eps) / invK0;
        invK1 = -sbz_exp(-tauHe) * invK0;
        T sq = sbz_sqrt(2.0 * PI / (beta * mHe * c_ * c_));
        T gcommon = 3.0 * fHe * (omXHe + eps) * c_ * c_ / (8.0 * PI * sq * (omXH + eps));
        T g2ps = gcommon * A2ps / (1.436289e-22 * fHe_2p_1s * fHe_2p_1s * fHe_2p_1s);
        invK2 = A2ps / (1.0 + 0.36 * sbz_pow(g2ps, 0.86)) * 3.0 * nHe * omXHe;
        T alphaHet = sbz_alphaHe_fit(Tb, pow(10.0, -16.306), 0.761);
        T betaHet = 4.0 / 3.0 * alphaHet / lam3 * sbz_exp(-beta * EHet_inf_2s);
        T tauHet = A2pt * nHe * (omXHe + eps) * 3.0 * (lamHet_2p_1s * lamHet_2p_1s * lamHet_2p_1s) /
                   (8.0 * PI * HSI);
        T pHet = (1.0 - sbz_exp(-tauHet)) / tauHet;
        T g2pt = gcommon * A2pt / (1.484872e-22 * fHet_2p_1s * fHet_2p_1s * fHet_2p_1s);
        T CHetnum = A2pt * (pHet + 1.0 / (1.0 + 0.66 * sbz_pow(g2pt, 0.9)) / 3.0) * sbz_exp(-beta * EHet_2p_2s);
        T CHet = (eps + CHetnum) / (eps + CHetnum + betaHet);
        DXHet = pref * CHet * (alphaHet * XHe * ne - betaHet * omXHe * 3.0 * sbz_exp(-beta * EHet_2s_1s));
    }
    T KHe = 1.0 / (invK0 + invK1 + invK2);
    T eHe = sbz_exp(-beta * EHe_2p_2s);
    T CHev = (eHe + KHe * LamHe * nHe * omXHe) / (eHe + KHe * (LamHe + betaHe) * nHe * omXHe);
    T CHe = sbz_smoothifelse(XHe - XlimC, CHev, T(1.0), 1e3);
    T DXHe = pref * CHe * (alphaHe * XHe * ne - betaHe * omXHe * sbz_exp(-beta * EHe_2s_1s));
    du[2] = DXHe + DXHet;

    du[0] = a * d.Hc;                                       // D(a) = ℋ a
    T DTg = -th.Tg * d.Hc;                                  // D(T0/a)
    du[3] = th.DT - DTg;
    du[4] = th.kappadot;
}
