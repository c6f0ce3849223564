// CPU ORACLE (test infrastructure only) — a plain restatement of SymBoltz.jl's perturbation stage.
//
// Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may load this library, and
// only as the checker.  The product path (libsbz_hip.so) never calls it.
//
// What it restates, function by function:
//   spline_eval      cubic Hermite background spline with extension extrapolation
//                    (src/utils.jl:118-127 → DataInterpolations.CubicHermiteSpline)
//   background_at    the observed background quantities the perturbation equations read:
//                    densities (generic_species.jl:23-24, neutrinos.jl:100-105, dark_energy.jl:47),
//                    ℋ (gravity.jl:31, metric.jl:26), X_e with He++ Saha and tanh reionization
//                    (baryons.jl:64-69,128-138,201-207), κ̇ (:175), T_b, DT, c_s² (:182-188)
//   rhs              the perturbation right-hand side: gravity.jl:38-39, photons.jl:30-42,56-59,
//                    neutrinos.jl:23-26 (massless) and :109-133 (massive), generic_species.jl:46-47,
//                    baryons (θ interaction cosmologies.jl:95), dark_energy.jl:50-58,
//                    couplings cosmologies.jl:85-97
//   jacobian         J(τ) = ∂f/∂u, built column by column as f(τ, e_j): exact because the
//                    perturbation system is linear in u (src/solve.jl:554)
//   dfdt             ∂f/∂τ by forward-mode dual numbers through rhs, as OrdinaryDiffEqDifferentiation
//                    does with ForwardDiff
//   initial_state    adiabatic initial conditions (cosmologies.jl:74-77, photons.jl:47-53,61-67,
//                    neutrinos.jl:33-39,134-139, generic_species.jl:54-57, dark_energy.jl:61-64)
//   solve_mode       Rodas5P (src/solve.jl:339) with a dense partial-pivoting LU in place of KLU
//                    (same solution of W x = b), RMS error norm, PI step control, Hairer–Wanner
//                    initial dt, saveat through cubic Hermite interpolation between steps
//   delta_m, pk      Δ_m = Σρ_sΔ_s/Σρ_s over c, b, h and P = P0 Δ_m² (fourier.jl:14-52,90-97)
//
// Parity pinning is done in tests/: golden CLASS P(k) of the reference (test/class_Pk.dat),
// self-convergence against rtol = 1e-10, J and ∂f/∂τ against finite differences, f = J u, the
// initial-condition invariants of test/runtests.jl:232-278, and the integrator's convergence
// order on problems with closed-form solutions.
#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <thread>
#include <set>
#include <vector>

namespace ora {

// ---------------------------------------------------------------- dual numbers (ForwardDiff)
struct D {
    double v, d;
    D(double x = 0.0, double dx = 0.0) : v(x), d(dx) {}
};
inline D operator+(D a, D b) { return {a.v + b.v, a.d + b.d}; }
inline D operator-(D a, D b) { return {a.v - b.v, a.d - b.d}; }
inline D operator-(D a) { return {-a.v, -a.d}; }
inline D operator*(D a, D b) { return {a.v * b.v, a.d * b.v + a.v * b.d}; }
inline D operator/(D a, D b) { return {a.v / b.v, (a.d * b.v - a.v * b.d) / (b.v * b.v)}; }
inline D operator+(D a, double b) { return {a.v + b, a.d}; }
inline D operator+(double a, D b) { return {a + b.v, b.d}; }
inline D operator-(D a, double b) { return {a.v - b, a.d}; }
inline D operator-(double a, D b) { return {a - b.v, -b.d}; }
inline D operator*(D a, double b) { return {a.v * b, a.d * b}; }
inline D operator*(double a, D b) { return {a * b.v, a * b.d}; }
inline D operator/(D a, double b) { return {a.v / b, a.d / b}; }
inline D operator/(double a, D b) { return {a / b.v, -a * b.d / (b.v * b.v)}; }
inline D exp(D x) { double e = std::exp(x.v); return {e, e * x.d}; }
inline D sqrt(D x) { double s = std::sqrt(x.v); return {s, x.d / (2 * s)}; }
inline D tanh(D x) { double t = std::tanh(x.v); return {t, (1 - t * t) * x.d}; }
inline D pow(D x, double p) { double y = std::pow(x.v, p); return {y, p * std::pow(x.v, p - 1) * x.d}; }
inline D fabs(D x) { return x.v < 0 ? -x : x; }
using std::exp;
using std::sqrt;
using std::tanh;
using std::pow;
using std::fabs;

// second-order Taylor numbers (value, first and second derivative), for v'' in the CMB sources
struct T2 {
    double v, d, dd;
    T2(double x = 0.0, double dx = 0.0, double ddx = 0.0) : v(x), d(dx), dd(ddx) {}
};
inline T2 operator+(T2 a, T2 b) { return {a.v + b.v, a.d + b.d, a.dd + b.dd}; }
inline T2 operator-(T2 a, T2 b) { return {a.v - b.v, a.d - b.d, a.dd - b.dd}; }
inline T2 operator-(T2 a) { return {-a.v, -a.d, -a.dd}; }
inline T2 operator*(T2 a, T2 b) { return {a.v * b.v, a.v * b.d + a.d * b.v, a.v * b.dd + 2 * a.d * b.d + a.dd * b.v}; }
inline T2 inv(T2 b) {  // 1/b: (1/b)' = -b'/b², (1/b)'' = 2b'²/b³ - b''/b²
    double r = 1.0 / b.v;
    return {r, -b.d * r * r, 2 * b.d * b.d * r * r * r - b.dd * r * r};
}
inline T2 operator/(T2 a, T2 b) { return a * inv(b); }
inline T2 operator+(T2 a, double b) { return {a.v + b, a.d, a.dd}; }
inline T2 operator+(double a, T2 b) { return {a + b.v, b.d, b.dd}; }
inline T2 operator-(T2 a, double b) { return {a.v - b, a.d, a.dd}; }
inline T2 operator-(double a, T2 b) { return {a - b.v, -b.d, -b.dd}; }
inline T2 operator*(T2 a, double b) { return {a.v * b, a.d * b, a.dd * b}; }
inline T2 operator*(double a, T2 b) { return {a * b.v, a * b.d, a * b.dd}; }
inline T2 operator/(T2 a, double b) { return {a.v / b, a.d / b, a.dd / b}; }
inline T2 operator/(double a, T2 b) { return a * inv(b); }
// chain rule for y = φ(x): y' = φ' x', y'' = φ'' x'² + φ' x''
inline T2 chain(T2 x, double f0, double f1, double f2) { return {f0, f1 * x.d, f2 * x.d * x.d + f1 * x.dd}; }
inline T2 exp(T2 x) { double e = std::exp(x.v); return chain(x, e, e, e); }
inline T2 sqrt(T2 x) { double r = std::sqrt(x.v); return chain(x, r, 0.5 / r, -0.25 / (r * x.v)); }
inline T2 tanh(T2 x) { double t = std::tanh(x.v); return chain(x, t, 1 - t * t, -2 * t * (1 - t * t)); }
inline T2 pow(T2 x, double p) {
    return chain(x, std::pow(x.v, p), p * std::pow(x.v, p - 1), p * (p - 1) * std::pow(x.v, p - 2));
}
inline T2 fabs(T2 x) { return x.v < 0 ? -x : x; }

// ---------------------------------------------------------------- constants (constants.jl:3-24)
const double C_LIGHT = 299792458.0;
const double H_PLANCK = 6.62607015e-34;
const double K_B = 1.380649e-23;
const double G_N = 6.6743e-11;
const double SIGMA_SB = 5.670374419184429e-8;
const double A_RAD = 4.0 / C_LIGHT * SIGMA_SB;
const double SIGMA_T = 6.6524587321e-29;
const double MPC = 3.0856775814913673e22;
const double H100 = 1e5 / MPC;
const double EV = 1.602176634e-19;
const double ME_KG = 9.1093837015e-31;
const double AMU = 1.6605390666e-27;
const double MH_KG = 1.008 * AMU;
const double MHE_KG = 4.002602 * AMU;
const double PI = 3.141592653589793;

// parameter slots (include/sbz_params.h, restated as a data format)
enum {
    P_H = 0, P_T0 = 1, P_OMG = 2, P_OMNU = 3, P_OMC = 4, P_OMB = 5, P_OMH = 6, P_OMDE = 7, P_YHE = 8,
    P_FHE = 9, P_NH = 10, P_Y0H = 12, P_IRHO0 = 13, P_ZRE1 = 14, P_DZRE1 = 15, P_NRE1 = 16,
    P_ZRE2 = 17, P_DZRE2 = 18, P_NRE2 = 19, P_W0 = 20, P_WA = 21, P_CS2X = 22, P_TAU0 = 23,
    P_KAPPA0 = 24, P_AS = 26, P_NS = 27, P_KPIVOT = 28, P_REION = 30, P_X = 32, P_W = 40,
    P_NX = 48, P_DE = 51
};

struct Model {
    int lg, lnu, nx, lh, de;  // lnu < 0: no massless ν; nx == 0: no massive ν; de: 0 Λ, 1 w0wa
    int iPhi, iF, iG, iN, iH, iC, iB, iX, n;
    void layout() {
        int i = 0;
        iPhi = i++;
        iF = i; i += lg + 1;
        iG = i; i += lg + 1;
        iN = i; if (lnu >= 0) i += lnu + 1;
        iH = i; if (nx > 0) i += nx * (lh + 1);
        iC = i; i += 2;
        iB = i; i += 2;
        iX = i; if (de == 1) i += 2;
        n = i;
    }
};

struct Spline {
    const double* t;
    const double* u;   // [nk][5]
    const double* du;  // [nk][5]
    int nk;
};

// Cubic Hermite spline of the five background unknowns with extension extrapolation.
template <class T>
void spline_eval(const Spline& s, T tau, T out[5]) {
    double tv = tau.v;
    int lo = 0, hi = s.nk - 1;
    if (tv <= s.t[0]) lo = 0;
    else if (tv >= s.t[s.nk - 1]) lo = s.nk - 2;
    else {
        while (hi - lo > 1) {
            int mid = (lo + hi) / 2;
            if (s.t[mid] <= tv) lo = mid; else hi = mid;
        }
    }
    double t0 = s.t[lo], t1 = s.t[lo + 1], h = t1 - t0;
    T x = (tau - t0) / h;
    T x2 = x * x, x3 = x2 * x;
    T h00 = 2.0 * x3 - 3.0 * x2 + 1.0, h10 = x3 - 2.0 * x2 + x, h01 = -2.0 * x3 + 3.0 * x2, h11 = x3 - x2;
    for (int j = 0; j < 5; ++j)
        out[j] = h00 * s.u[lo * 5 + j] + h10 * (h * s.du[lo * 5 + j]) + h01 * s.u[(lo + 1) * 5 + j] +
                 h11 * (h * s.du[(lo + 1) * 5 + j]);
}
inline void spline_eval(const Spline& s, double tau, double out[5]) {
    D o[5];
    spline_eval<D>(s, D(tau, 0.0), o);
    for (int j = 0; j < 5; ++j) out[j] = o[j].v;
}

template <class T>
struct Bg {
    T a, H, rg, rn, rc, rb, rh, ph, Irho, IP, rx, wx, wdotx, kd, cs2, kappa, Xe, Tb;
    T E[8];
};

template <class T>
T smooth_step(T x, double v1, double v2, double k) { return 0.5 * ((v1 + v2) + (v2 - v1) * tanh(k * x)); }

template <class T>
T reion(T a, double zre, double dz, double n, double xmax) {
    T z1 = 1.0 / a;
    T fz = (n == 1.5) ? sqrt(z1) * z1 : pow(z1, n);
    double fr = std::pow(1.0 + zre, n);
    return smooth_step(fr - fz, 0.0, xmax, 1.0 / (n * std::pow(1.0 + zre, n - 1.0) * dz));
}

template <class T>
void background_at(const double* P, const Spline& s, T tau, Bg<T>& b) {
    T sv[5];
    spline_eval(s, tau, sv);
    T a = sv[0], XH = sv[1], XHe = sv[2], dT = sv[3], kap = sv[4];
    b.a = a;
    const double q = 3.0 / (8.0 * PI);
    T a2 = a * a;
    b.rg = q * P[P_OMG] / (a2 * a2);
    b.rn = q * P[P_OMNU] / (a2 * a2);
    b.rc = q * P[P_OMC] / (a2 * a);
    b.rb = q * P[P_OMB] / (a2 * a);
    int nx = (int)P[P_NX];
    if (P[P_NH] > 0) {
        T y = P[P_Y0H] * a;
        T Ir = 0.0, Ip = 0.0;
        for (int i = 0; i < nx; ++i) {
            double x = P[P_X + i];
            b.E[i] = sqrt(x * x + y * y);
            Ir = Ir + P[P_W + i] * b.E[i];
            Ip = Ip + P[P_W + i] * x * x / b.E[i];
        }
        b.Irho = Ir;
        b.IP = Ip;
        b.rh = q * P[P_OMH] * Ir / (P[P_IRHO0] * a2 * a2);
        b.ph = q * P[P_OMH] * Ip / (3.0 * P[P_IRHO0] * a2 * a2);
    } else {
        b.rh = b.ph = 0.0;
        b.Irho = 1.0;
        b.IP = 0.0;
    }
    if (P[P_DE] == 1.0) {
        double w0 = P[P_W0], wa = P[P_WA];
        b.rx = q * P[P_OMDE] * pow(fabs(a), -3.0 * (1.0 + w0 + wa)) * exp(-3.0 * wa * (1.0 - a));
        b.wx = w0 + wa * (1.0 - a);
    } else {
        b.rx = q * P[P_OMDE];
        b.wx = -1.0;
    }
    T rho = b.rg + b.rn + b.rc + b.rb + b.rh + b.rx;
    b.H = sqrt(8.0 * PI / 3.0 * rho) * a;
    b.wdotx = (P[P_DE] == 1.0) ? -P[P_WA] * a * b.H : T(0.0);  // D(w) = -wa D(a), D(a) = a ℋ
    // thermodynamics
    double h = P[P_H], YHe = P[P_YHE], fHe = P[P_FHE];
    double H0 = H100 * h;
    T Tg = P[P_T0] / a;
    T Tb = dT + Tg;
    T nH = (1.0 - YHe) * b.rb * H0 * H0 / G_N / MH_KG;
    T beta = 1.0 / (K_B * Tb);
    T lam = H_PLANCK / sqrt(2.0 * PI * ME_KG / beta);
    T R = exp(-beta * (54.4178 * EV)) / (nH * lam * lam * lam);
    T dd = 1.0 + fHe + R;
    T Xpp = 2.0 * R * fHe / dd / (1.0 + sqrt(1.0 + 4.0 * R * fHe / (dd * dd)));
    T Xe = XH + fHe * XHe + Xpp;
    if (P[P_REION] != 0) {
        Xe = Xe + reion(a, P[P_ZRE1], P[P_DZRE1], P[P_NRE1], 1.0 + fHe);
        Xe = Xe + reion(a, P[P_ZRE2], P[P_DZRE2], P[P_NRE2], fHe);
    }
    b.Xe = Xe;
    b.Tb = Tb;
    b.kd = -a / H0 * (Xe * nH) * SIGMA_T * C_LIGHT;
    T DT = -2.0 * Tb * b.H - a / h * 8.0 / 3.0 * SIGMA_T * A_RAD / H100 * (Tg * Tg * Tg * Tg) / (ME_KG * C_LIGHT) *
           Xe / (1.0 + fHe + Xe) * dT;
    T mu = MH_KG * C_LIGHT * C_LIGHT / (1.0 + (MH_KG / MHE_KG - 1.0) * YHe + Xe * (1.0 - YHe));
    b.cs2 = K_B / mu * (Tb - DT / (3.0 * b.H));
    b.kappa = kap - P[P_KAPPA0];
}

inline double dlnf0(double x) { return -x / (1.0 + std::exp(-x)); }

// Perturbation right-hand side (linear in u).
template <class T, class U>
void rhs(const Model& m, const double* P, const Bg<T>& b, double k, T tau, const U* u, U* du) {
    const double* X = P + P_X;
    const double* W = P + P_W;
    int nx = m.nx;
    const U* F = u + m.iF;
    const U* G = u + m.iG;
    U* dF = du + m.iF;
    U* dG = du + m.iG;
    T a2 = b.a * b.a;
    // massive ν moments
    U Idr = 0.0, sig_h_num = 0.0;
    if (nx > 0) {
        for (int i = 0; i < nx; ++i) {
            const U* ps = u + m.iH + i * (m.lh + 1);
            Idr = Idr + W[i] * b.E[i] * ps[0];
            sig_h_num = sig_h_num + W[i] * X[i] * X[i] / b.E[i] * ps[2];
        }
    }
    // anisotropic stress Π = Σ (1+w) ρ σ and density perturbation δρ = Σ δ ρ
    U Pi = (4.0 / 3.0) * b.rg * (F[2] / 2.0);
    U drho = F[0] * b.rg + u[m.iC] * b.rc + u[m.iB] * b.rb;
    if (m.lnu >= 0) {
        const U* N = u + m.iN;
        Pi = Pi + (4.0 / 3.0) * b.rn * (N[2] / 2.0);
        drho = drho + N[0] * b.rn;
    }
    if (nx > 0) {
        U sig_h = (2.0 / 3.0) * sig_h_num / (b.Irho + b.IP / 3.0);
        Pi = Pi + (b.rh + b.ph) * sig_h;
        drho = drho + (Idr / b.Irho) * b.rh;
    }
    if (m.de == 1) drho = drho + u[m.iX] * b.rx;
    U Phi = u[m.iPhi];
    U Psi = Phi - 12.0 * PI * a2 * Pi / (k * k);
    U dPhi = -4.0 * PI / 3.0 * a2 / b.H * drho - k * k / (3.0 * b.H) * Phi - b.H * Psi;
    du[m.iPhi] = dPhi;
    // photons
    int L = m.lg;
    U thg = 3.0 * k * F[1] / 4.0;
    U thb = u[m.iB + 1];
    U Pig = F[2] + G[0] + G[2];
    dF[0] = -k * F[1] + 4.0 * dPhi;
    dF[1] = k / 3.0 * (F[0] - 2.0 * F[2] + 4.0 * Psi) - 4.0 / 3.0 * b.kd / k * (thb - thg);
    for (int l = 2; l < L; ++l)
        dF[l] = k / (2 * l + 1) * (l * F[l - 1] - (l + 1) * F[l + 1]) + b.kd * (F[l] - (l == 2 ? Pig / 10.0 : U(0.0)));
    dF[L] = k * F[L - 1] - (L + 1) / tau * F[L] + b.kd * F[L];
    dG[0] = k * (-G[1]) + b.kd * (G[0] - Pig / 2.0);
    dG[1] = k / 3.0 * (G[0] - 2.0 * G[2]) + b.kd * G[1];
    for (int l = 2; l < L; ++l)
        dG[l] = k / (2 * l + 1) * (l * G[l - 1] - (l + 1) * G[l + 1]) + b.kd * (G[l] - (l == 2 ? Pig / 10.0 : U(0.0)));
    dG[L] = k * G[L - 1] - (L + 1) / tau * G[L] + b.kd * G[L];
    // massless neutrinos
    if (m.lnu >= 0) {
        const U* N = u + m.iN;
        U* dN = du + m.iN;
        int Ln = m.lnu;
        dN[0] = -k * N[1] + 4.0 * dPhi;
        dN[1] = k / 3.0 * (N[0] - 2.0 * N[2] + 4.0 * Psi);
        for (int l = 2; l < Ln; ++l) dN[l] = k / (2 * l + 1) * (l * N[l - 1] - (l + 1) * N[l + 1]);
        dN[Ln] = k * N[Ln - 1] - (Ln + 1) / tau * N[Ln];
    }
    // massive neutrinos
    for (int i = 0; i < nx; ++i) {
        const U* ps = u + m.iH + i * (m.lh + 1);
        U* dps = du + m.iH + i * (m.lh + 1);
        int Lh = m.lh;
        T xe = X[i] / b.E[i];
        double df = dlnf0(X[i]);
        dps[0] = -k * xe * ps[1] - dPhi * df;
        dps[1] = k / 3.0 * xe * (ps[0] - 2.0 * ps[2]) - k / 3.0 * (b.E[i] / X[i]) * Psi * df;
        for (int l = 2; l < Lh; ++l) dps[l] = k / (2 * l + 1) * xe * (l * ps[l - 1] - (l + 1) * ps[l + 1]);
        dps[Lh] = k / (2 * Lh + 1) * xe *
                  (Lh * ps[Lh - 1] - (Lh + 1) * ((2 * Lh + 1) * (b.E[i] / X[i]) * ps[Lh] / (k * tau) - ps[Lh - 1]));
    }
    // cold dark matter (w = c_s² = 0) and baryons
    U dc = u[m.iC], thc = u[m.iC + 1], db = u[m.iB];
    du[m.iC] = -(thc - 3.0 * dPhi);
    du[m.iC + 1] = -b.H * thc + k * k * Psi;
    du[m.iB] = -(thb - 3.0 * dPhi) - 3.0 * b.H * b.cs2 * db;
    U thint = -b.kd * 4.0 * b.rg / (3.0 * b.rb) * (thg - thb);
    du[m.iB + 1] = -b.H * thb + b.cs2 * k * k * db + k * k * Psi + thint;
    (void)dc;
    // w0wa fluid
    if (m.de == 1) {
        U dx = u[m.iX], thx = u[m.iX + 1];
        T w = b.wx;
        double cs2 = P[P_CS2X];
        T ca2 = w - b.wdotx / (3.0 * b.H * (1.0 + w));
        du[m.iX] = -(1.0 + w) * (thx - 3.0 * dPhi) - 3.0 * b.H * (cs2 - w) * dx -
                   9.0 * (b.H / k) * (b.H / k) * (1.0 + w) * (cs2 - ca2) * thx;
        du[m.iX + 1] = -b.H * (1.0 - 3.0 * cs2) * thx + cs2 / (1.0 + w) * k * k * dx + k * k * Psi;
    }
}

void f_eval(const Model& m, const double* P, const Spline& s, double k, double tau, const double* u, double* du) {
    Bg<double> b;
    background_at<double>(P, s, tau, b);
    rhs<double, double>(m, P, b, k, tau, u, du);
}

// J(τ) column by column: J e_j = f(τ, e_j) for the linear system.
void jacobian(const Model& m, const double* P, const Spline& s, double k, double tau, double* J) {
    int n = m.n;
    Bg<double> b;
    background_at<double>(P, s, tau, b);
    std::vector<double> e(n, 0.0), col(n);
    for (int j = 0; j < n; ++j) {
        e[j] = 1.0;
        rhs<double, double>(m, P, b, k, tau, e.data(), col.data());
        for (int i = 0; i < n; ++i) J[i * n + j] = col[i];
        e[j] = 0.0;
    }
}

// ∂f/∂τ at fixed u by dual numbers through the background and the rhs.
void dfdt(const Model& m, const double* P, const Spline& s, double k, double tau, const double* u, double* out) {
    int n = m.n;
    Bg<D> b;
    D t(tau, 1.0);
    background_at<D>(P, s, t, b);
    std::vector<D> ud(n), dd(n);
    for (int i = 0; i < n; ++i) ud[i] = D(u[i], 0.0);
    rhs<D, D>(m, P, b, k, t, ud.data(), dd.data());
    for (int i = 0; i < n; ++i) out[i] = dd[i].d;
}

void initial_state(const Model& m, const double* P, const Spline& s, double k, double tau, double* u) {
    int n = m.n;
    std::fill(u, u + n, 0.0);
    Bg<double> b;
    background_at<double>(P, s, tau, b);
    double fnu = (b.rn + b.rh) / (b.rg + b.rn + b.rh);
    double Psi = 20.0 * 0.5 / (15.0 + 4.0 * fnu);
    double kt = k * tau;
    double* F = u + m.iF;
    double* G = u + m.iG;
    F[0] = -2.0 * Psi;
    F[1] = 2.0 / 3.0 * kt * Psi;
    F[2] = -8.0 / 15.0 * k / b.kd * F[1];
    F[3] = -3.0 / 7.0 * k / b.kd * F[2];
    G[0] = 5.0 / 16.0 * F[2];
    G[1] = -1.0 / 16.0 * k / b.kd * F[2];
    G[2] = 1.0 / 16.0 * F[2];
    G[3] = -3.0 / 7.0 * k / b.kd * G[2];
    if (m.lnu >= 0) {
        double* N = u + m.iN;
        N[0] = -2.0 * Psi;
        N[1] = 4.0 / (3.0 * k) * (0.5 * k * k * tau * Psi);
        N[2] = 2.0 * (1.0 / 15.0 * kt * kt * Psi);
        N[3] = 3.0 / 7.0 * kt * N[2];
    }
    for (int i = 0; i < m.nx; ++i) {
        double* ps = u + m.iH + i * (m.lh + 1);
        double x = P[P_X + i], df = dlnf0(x);
        ps[0] = -0.25 * (-2.0 * Psi) * df;
        ps[1] = -1.0 / 3.0 * b.E[i] / x * (0.5 * kt * Psi) * df;
        ps[2] = -0.5 * (1.0 / 15.0 * kt * kt * Psi) * df;
    }
    u[m.iC] = -1.5 * Psi;
    u[m.iC + 1] = 0.5 * k * k * tau * Psi;
    u[m.iB] = -1.5 * Psi;
    u[m.iB + 1] = 0.5 * k * k * tau * Psi;
    if (m.de == 1) {
        u[m.iX] = -1.5 * (1.0 + b.wx) * Psi;
        u[m.iX + 1] = 0.5 * k * k * tau * Psi;
    }
    // Φ from the constraint k²(Φ - Ψ) = 12π a² Π (gravity.jl:39) with Π of the state above:
    // evaluate Ψ(Φ=0) from the rhs machinery's definition.
    u[m.iPhi] = 0.0;
    double Pi = (4.0 / 3.0) * b.rg * (F[2] / 2.0);
    if (m.lnu >= 0) Pi += (4.0 / 3.0) * b.rn * (u[m.iN + 2] / 2.0);
    if (m.nx > 0) {
        double num = 0.0;
        for (int i = 0; i < m.nx; ++i) {
            double x = P[P_X + i];
            num += P[P_W + i] * x * x / b.E[i] * u[m.iH + i * (m.lh + 1) + 2];
        }
        Pi += (b.rh + b.ph) * (2.0 / 3.0) * num / (b.Irho + b.IP / 3.0);
    }
    u[m.iPhi] = Psi + 12.0 * PI * b.a * b.a * Pi / (k * k);
}

// Gauge-invariant matter overdensity Δ_m = Σ ρ_s Δ_s / Σ ρ_s over c, b, h (fourier.jl:39-52).
double delta_m(const Model& m, const double* P, const Spline& s, double k, double tau, const double* u) {
    Bg<double> b;
    background_at<double>(P, s, tau, b);
    double H = b.H;
    double Dc = u[m.iC] + 3.0 * H * u[m.iC + 1] / (k * k);
    double Db = u[m.iB] + 3.0 * H * u[m.iB + 1] / (k * k);
    double num = b.rc * Dc + b.rb * Db, den = b.rc + b.rb;
    if (m.nx > 0) {
        double Idr = 0.0, uh = 0.0;
        for (int i = 0; i < m.nx; ++i) {
            const double* ps = u + m.iH + i * (m.lh + 1);
            Idr += P[P_W + i] * b.E[i] * ps[0];
            uh += P[P_W + i] * P[P_X + i] * ps[1];
        }
        double dh = Idr / b.Irho;
        double th = uh / (b.Irho + b.IP / 3.0) * k;
        double wh = b.ph / b.rh;
        double Dh = dh + 3.0 * H * (1.0 + wh) * th / (k * k);
        num += b.rh * Dh;
        den += b.rh;
    }
    return num / den;
}

double p0(const double* P, double k) {
    return 2.0 * PI * PI / (k * k * k) * P[P_AS] * std::pow(k / P[P_KPIVOT], P[P_NS] - 1.0);
}

// Ψ = Φ - 12π a² Π / k² for a state (linear in u), Π = Σ (1+w) ρ σ (gravity.jl:39, cosmologies.jl:94).
template <class T, class U>
U psi_of(const Model& m, const double* P, const Bg<T>& b, double k, const U* u) {
    const U* F = u + m.iF;
    U Pi = (4.0 / 3.0) * b.rg * (F[2] / 2.0);
    if (m.lnu >= 0) Pi = Pi + (4.0 / 3.0) * b.rn * (u[m.iN + 2] / 2.0);
    if (m.nx > 0) {
        U num = 0.0;
        for (int i = 0; i < m.nx; ++i) {
            double x = P[P_X + i];
            num = num + P[P_W + i] * x * x / b.E[i] * u[m.iH + i * (m.lh + 1) + 2];
        }
        Pi = Pi + (b.rh + b.ph) * ((2.0 / 3.0) * num / (b.Irho + b.IP / 3.0));
    }
    return u[m.iPhi] - 12.0 * PI * b.a * b.a * Pi / (k * k);
}

// Visibility v = D(e^{-κ}) = -κ̇ e^{-κ} (baryons.jl:179) and its first two τ-derivatives
// v' = (-κ̈ + κ̇²) e^{-κ}, v'' = (-κ⃛ + 3 κ̇ κ̈ - κ̇³) e^{-κ}, with κ̇ and its derivatives from
// second-order Taylor numbers through the background; out = (v, v', v'', e^{-κ}).
void visibility(const double* P, const Spline& s, double tau, double* out) {
    Bg<T2> bt;
    background_at<T2>(P, s, T2(tau, 1.0, 0.0), bt);
    double kd = bt.kd.v, kd1 = bt.kd.d, kd2 = bt.kd.dd, eK = std::exp(-bt.kappa.v);
    out[0] = -kd * eK;
    out[1] = (-kd1 + kd * kd) * eK;
    out[2] = (-kd2 + 3 * kd * kd1 - kd * kd * kd) * eK;
    out[3] = eK;
}

// CMB source functions at (τ, u) (cosmologies.jl:99-105):
//   S_T = v(δγ/4 + Ψ + Πγ/16) + e^{-κ}(Ψ' + Φ') + (v θ_b)'/k² + 3/(16k²)(v Πγ)'',
//   S_E = 3 v Πγ / (16 (kχ)²),   S_ψ = -(Ψ+Φ)(τ-τrec)/(τ0-τrec)/(τ0-τ) for τ ≥ τrec.
// u' = f(τ,u), u'' = ∂f/∂τ + J u'; v'' through second-order Taylor numbers of the background.
void sources(const Model& m, const double* P, const Spline& s, double k, double tau, const double* u, double* S) {
    int n = m.n;
    std::vector<double> f(n), g(n), tmp(n);
    f_eval(m, P, s, k, tau, u, f.data());
    f_eval(m, P, s, k, tau, f.data(), g.data());
    dfdt(m, P, s, k, tau, u, tmp.data());
    for (int i = 0; i < n; ++i) g[i] += tmp[i];
    double vis[4];
    visibility(P, s, tau, vis);
    const double v = vis[0], vd = vis[1], vdd = vis[2], eK = vis[3];
    Bg<double> b;
    background_at<double>(P, s, tau, b);
    Bg<D> bd;
    background_at<D>(P, s, D(tau, 1.0), bd);
    std::vector<D> ud(u, u + n);
    double Psi = psi_of<double, double>(m, P, b, k, u);
    double Psid = psi_of<D, D>(m, P, bd, k, ud.data()).d + psi_of<double, double>(m, P, b, k, f.data());
    const double* F = u + m.iF;
    const double* G = u + m.iG;
    double Pi = F[2] + G[0] + G[2];
    double Pid = f[m.iF + 2] + f[m.iG] + f[m.iG + 2];
    double Pidd = g[m.iF + 2] + g[m.iG] + g[m.iG + 2];
    double thb = u[m.iB + 1], thbd = f[m.iB + 1];
    double tau0 = P[P_TAU0], taurec = P[25];
    double ST = v * (F[0] / 4.0 + Psi + Pi / 16.0) + eK * (Psid + f[m.iPhi]) + (vd * thb + v * thbd) / (k * k) +
                3.0 / (16.0 * k * k) * (vdd * Pi + 2.0 * vd * Pid + v * Pidd);
    double chi = tau0 - tau;
    double SE = chi > 0 ? 3.0 / 16.0 * v * Pi / ((k * chi) * (k * chi)) : 0.0;
    double Spsi = (tau >= taurec && chi > 0) ? -(Psi + u[m.iPhi]) * (tau - taurec) / (tau0 - taurec) / chi : 0.0;
    S[0] = ST;
    S[1] = SE;
    S[2] = Spsi;
}

// ---------------------------------------------------------------- dense LU (partial pivoting)
bool lu_factor(int n, double* A, int* piv) {
    for (int c = 0; c < n; ++c) {
        int p = c;
        double mx = std::fabs(A[c * n + c]);
        for (int i = c + 1; i < n; ++i) {
            double v = std::fabs(A[i * n + c]);
            if (v > mx) { mx = v; p = i; }
        }
        if (!(mx > 0.0) || !std::isfinite(mx)) return false;
        piv[c] = p;
        if (p != c)
            for (int j = 0; j < n; ++j) std::swap(A[c * n + j], A[p * n + j]);
        double r = 1.0 / A[c * n + c];
        for (int i = c + 1; i < n; ++i) {
            double l = A[i * n + c] * r;
            A[i * n + c] = l;
            if (l != 0.0)
                for (int j = c + 1; j < n; ++j) A[i * n + j] -= l * A[c * n + j];
        }
    }
    return true;
}
// Full-row swaps during factorization: apply every interchange to b first, then substitute.
void lu_solve(int n, const double* A, const int* piv, double* b) {
    for (int c = 0; c < n; ++c)
        if (piv[c] != c) std::swap(b[c], b[piv[c]]);
    for (int i = 1; i < n; ++i) {
        double s = b[i];
        for (int j = 0; j < i; ++j) s -= A[i * n + j] * b[j];
        b[i] = s;
    }
    for (int i = n - 1; i >= 0; --i) {
        double s = b[i];
        for (int j = i + 1; j < n; ++j) s -= A[i * n + j] * b[j];
        b[i] = s / A[i * n + i];
    }
}

// ---------------------------------------------------------------- Rodas5P
const double GAM = 0.21193756319429014;
const double RA[9][8] = {
    {0}, {0}, {3.0}, {2.849394379747939, 0.45842242204463923},
    {-6.954028509809101, 2.489845061869568, -10.358996098473584},
    {2.8029986275628964, 0.5072464736228206, -0.3988312541770524, -0.04721187230404641},
    {-7.502846399306121, 2.561846144803919, -11.627539656261098, -0.18268767659942256, 0.030198172008377946},
    {-7.502846399306121, 2.561846144803919, -11.627539656261098, -0.18268767659942256, 0.030198172008377946, 1.0},
    {-7.502846399306121, 2.561846144803919, -11.627539656261098, -0.18268767659942256, 0.030198172008377946, 1.0, 1.0}};
const double RC[9][8] = {
    {0}, {0}, {-14.155112264123755}, {-17.97296035885952, -2.859693295451294},
    {147.12150275711716, -1.41221402718213, 71.68940251302358},
    {165.43517024871676, -0.4592823456491126, 42.90938336958603, -5.961986721573306},
    {24.854864614690072, -3.0009227002832186, 47.4931110020768, 5.5814197821558125, -0.6610691825249471},
    {30.91273214028599, -3.1208243349937974, 77.79954646070892, 34.28646028294783, -19.097331116725623, -28.087943162872662},
    {37.80277123390563, -3.2571969029072276, 112.26918849496327, 66.9347231244047, -40.06618937091002, -54.66780262877968, -9.48861652309627}};
const double RCC[9] = {0, 0, 0.6358126895828704, 0.4095798393397535, 0.9769306725060716, 0.4288403609558664, 1, 1, 1};
const double RD[9] = {0, 0.21193756319429014, -0.42387512638858027, -0.3384627126235924, 1.8046452872882734,
                      2.325825639765069, 0, 0, 0};

struct Opts {
    double rtol, atol;
    long maxiters;
    bool sparse = false;   // J by column coloring + static-pattern sparse LU (the CPU baseline path)
};

struct Stats {
    int64_t naccept, nreject, nf, njac, nlu, nsolve;
};

enum { RET_SUCCESS = 0, RET_MAXITERS = 1, RET_DTMIN = 2, RET_UNSTABLE = 3, RET_INITFAIL = 4 };

double rms_scaled(int n, const double* x, const double* sk) {
    double s = 0.0;
    for (int i = 0; i < n; ++i) { double r = x[i] / sk[i]; s += r * r; }
    return std::sqrt(s / n);
}

// Integrate one k-mode from t0 to t1.  saveat: sorted times in [t0, t1] at which the state is
// stored by cubic Hermite interpolation over the accepted step that contains them.
// ---------------------------------------------------------------- sparse path
// The structure of the reference's default solver: an analytic sparse Jacobian (MTK `jac = true,
// sparse = true`, src/solve.jl:126-127) factored by KLU (src/solve.jl:327-343).  Here: the pattern
// of J is found once per mode; J is then assembled per step from one background evaluation and
// one RHS application per column color (f is linear in u, so f(Σ_{j∈c} e_j) holds column j of J
// in every row that only column j of color c reaches); W = I/(γh) - J is factored on its static
// fill pattern in a minimum-degree elimination order, without pivoting, like KLU's numeric
// refactorization on a fixed symbolic analysis.
struct Sparse {
    int n = 0, ncolor = 0;
    std::vector<std::vector<int>> col_rows;    // rows of the J pattern in each column
    std::vector<std::vector<int>> color_cols;  // columns of each color
    std::vector<int> order;                    // elimination order (pivots)
    std::vector<std::vector<int>> lower, upper;  // per elimination step: rows below, columns right
};

Sparse sparse_structure(const Model& m, const double* P, const Spline& s, double k, double t0, double t1) {
    const int n = m.n;
    Sparse sp;
    sp.n = n;
    std::vector<char> pat((size_t)n * n, 0);
    std::vector<double> J((size_t)n * n);
    const double ts[3] = {t0, std::sqrt(std::max(t0, 1e-300) * t1), t1};
    for (double tau : ts) {
        jacobian(m, P, s, k, tau, J.data());
        for (size_t q = 0; q < J.size(); ++q)
            if (J[q] != 0.0) pat[q] = 1;
    }
    sp.col_rows.assign(n, {});
    std::vector<std::vector<int>> row_cols(n);
    for (int i = 0; i < n; ++i)
        for (int j = 0; j < n; ++j)
            if (pat[(size_t)i * n + j]) {
                sp.col_rows[j].push_back(i);
                row_cols[i].push_back(j);
            }
    // greedy coloring: columns that share a row get different colors
    std::vector<int> color(n, -1);
    for (int j = 0; j < n; ++j) {
        std::vector<char> used(n + 1, 0);
        for (int i : sp.col_rows[j])
            for (int c2 : row_cols[i])
                if (color[c2] >= 0) used[color[c2]] = 1;
        int c = 0;
        while (used[c]) ++c;
        color[j] = c;
        sp.ncolor = std::max(sp.ncolor, c + 1);
    }
    sp.color_cols.assign(sp.ncolor, {});
    for (int j = 0; j < n; ++j) sp.color_cols[color[j]].push_back(j);
    // minimum-degree order on the symmetrized pattern, then the fill of elimination in that order
    std::vector<std::set<int>> adj(n);
    for (int i = 0; i < n; ++i)
        for (int j : row_cols[i])
            if (i != j) { adj[i].insert(j); adj[j].insert(i); }
    std::vector<char> alive(n, 1);
    for (int r = 0; r < n; ++r) {
        int best = -1;
        for (int v = 0; v < n; ++v)
            if (alive[v] && (best < 0 || adj[v].size() < adj[best].size())) best = v;
        sp.order.push_back(best);
        alive[best] = 0;
        for (int a : adj[best]) {
            adj[a].erase(best);
            for (int b : adj[best])
                if (b != a) adj[a].insert(b);
        }
        adj[best].clear();
    }
    std::vector<int> pos(n);
    for (int r = 0; r < n; ++r) pos[sp.order[r]] = r;
    std::vector<std::set<int>> rows(n), cols(n);   // filled pattern by row and by column (W = diag + J)
    for (int i = 0; i < n; ++i) {
        rows[i].insert(i);
        cols[i].insert(i);
        for (int j : row_cols[i]) { rows[i].insert(j); cols[j].insert(i); }
    }
    for (int r = 0; r < n; ++r) {
        const int p = sp.order[r];
        std::vector<int> lo, up;
        for (int i : cols[p]) if (pos[i] > r) lo.push_back(i);
        for (int j : rows[p]) if (pos[j] > r) up.push_back(j);
        for (int i : lo)
            for (int j : up)
                if (!rows[i].count(j)) { rows[i].insert(j); cols[j].insert(i); }
        sp.lower.push_back(lo);
        sp.upper.push_back(up);
    }
    return sp;
}

// J on the pattern at τ: one background evaluation, one RHS application per color.
void jacobian_colored(const Model& m, const double* P, const Spline& s, const Sparse& sp, double k, double tau,
                      double* J, double* e, double* col) {
    const int n = m.n;
    Bg<double> b;
    background_at<double>(P, s, tau, b);
    for (int c = 0; c < sp.ncolor; ++c) {
        for (int j : sp.color_cols[c]) e[j] = 1.0;
        rhs<double, double>(m, P, b, k, tau, e, col);
        for (int j : sp.color_cols[c]) {
            e[j] = 0.0;
            for (int i : sp.col_rows[j]) J[(size_t)i * n + j] = col[i];
        }
    }
}

// W = I/(γh) - J factored in place on its fill pattern; diagonal entries hold the reciprocal pivots.
bool sparse_lu(const Sparse& sp, double* W) {
    const int n = sp.n;
    for (int r = 0; r < n; ++r) {
        const int p = sp.order[r];
        const double piv = W[(size_t)p * n + p];
        if (!(piv != 0.0) || !std::isfinite(piv)) return false;
        const double rp = 1.0 / piv;
        W[(size_t)p * n + p] = rp;
        for (int i : sp.lower[r]) {
            const double l = W[(size_t)i * n + p] * rp;
            W[(size_t)i * n + p] = l;
            for (int j : sp.upper[r]) W[(size_t)i * n + j] -= l * W[(size_t)p * n + j];
        }
    }
    return true;
}
void sparse_solve(const Sparse& sp, const double* W, double* b) {
    const int n = sp.n;
    for (int r = 0; r < n; ++r) {
        const int p = sp.order[r];
        for (int i : sp.lower[r]) b[i] -= W[(size_t)i * n + p] * b[p];
    }
    for (int r = n - 1; r >= 0; --r) {
        const int p = sp.order[r];
        double acc = b[p];
        for (int j : sp.upper[r]) acc -= W[(size_t)p * n + j] * b[j];
        b[p] = acc * W[(size_t)p * n + p];
    }
}

int solve_mode(const Model& m, const double* P, const Spline& s, double k, double t0, double t1, const Opts& o,
               const double* saveat, int nsave, double* usave, double* ufinal, Stats& st) {
    int n = m.n;
    std::memset(&st, 0, sizeof(st));
    std::vector<double> u(n), u1(n), f0(n), f1(n), dT(n), J(n * n), W(n * n), us(n), rhsv(n), sk(n);
    std::vector<double> K(9 * n);
    std::vector<int> piv(n);
    if (!std::isfinite(k) || k <= 0.0) {
        for (int i = 0; i < n; ++i) ufinal[i] = NAN;
        return RET_INITFAIL;
    }
    initial_state(m, P, s, k, t0, u.data());
    for (int i = 0; i < n; ++i)
        if (!std::isfinite(u[i])) { std::copy(u.begin(), u.end(), ufinal); return RET_INITFAIL; }
    double t = t0;
    int isave = 0;
    while (isave < nsave && saveat[isave] <= t0) {
        std::copy(u.begin(), u.end(), usave + (size_t)isave * n);
        ++isave;
    }
    if (t1 <= t0) {
        std::copy(u.begin(), u.end(), ufinal);
        return RET_SUCCESS;
    }
    Sparse sp;
    std::vector<double> se, scol;
    if (o.sparse) {
        sp = sparse_structure(m, P, s, k, t0, t1);
        se.assign(n, 0.0);
        scol.assign(n, 0.0);
        std::fill(J.begin(), J.end(), 0.0);
    }
    // initial step (Hairer–Wanner, as OrdinaryDiffEqCore)
    f_eval(m, P, s, k, t, u.data(), f0.data());
    st.nf++;
    for (int i = 0; i < n; ++i) sk[i] = o.atol + std::fabs(u[i]) * o.rtol;
    double d0 = rms_scaled(n, u.data(), sk.data()), d1 = rms_scaled(n, f0.data(), sk.data());
    double dt0 = (d0 < 1e-5 || d1 < 1e-5) ? 1e-6 : 0.01 * d0 / d1;
    dt0 = std::min(dt0, t1 - t);
    for (int i = 0; i < n; ++i) u1[i] = u[i] + dt0 * f0[i];
    f_eval(m, P, s, k, t + dt0, u1.data(), f1.data());
    st.nf++;
    for (int i = 0; i < n; ++i) rhsv[i] = f1[i] - f0[i];
    double d2 = rms_scaled(n, rhsv.data(), sk.data()) / dt0;
    double mx = std::max(d1, d2);
    double dt1 = (mx <= 1e-15) ? std::max(1e-6, dt0 * 1e-3) : std::pow(0.01 / mx, 1.0 / 6.0);
    double dt = std::min(std::min(100.0 * dt0, dt1), t1 - t);
    double qold = 1e-4;
    long iters = 0;
    bool have_f0 = true;  // f0 holds f(t, u)
    while (t < t1) {
        if (++iters > o.maxiters) { std::copy(u.begin(), u.end(), ufinal); return RET_MAXITERS; }
        if (t + dt >= t1 - 1e-12 * std::fabs(t1)) dt = t1 - t;  // final step lands on t1
        else if (dt < 1e-14 * std::fabs(t)) { std::copy(u.begin(), u.end(), ufinal); return RET_DTMIN; }
        // Jacobian, time derivative and W = I/(γh) - J at (t, u)
        if (o.sparse) jacobian_colored(m, P, s, sp, k, t, J.data(), se.data(), scol.data());
        else jacobian(m, P, s, k, t, J.data());
        st.njac++;
        dfdt(m, P, s, k, t, u.data(), dT.data());
        if (!have_f0) { f_eval(m, P, s, k, t, u.data(), f0.data()); st.nf++; have_f0 = true; }
        double ig = 1.0 / (GAM * dt);
        bool ok;
        if (o.sparse) {   // W on the fill pattern (fill entries start at zero: J holds zeros there)
            for (int r = 0; r < n; ++r) {
                const int p = sp.order[r];
                W[(size_t)p * n + p] = ig - J[(size_t)p * n + p];
                for (int i : sp.lower[r]) W[(size_t)i * n + p] = -J[(size_t)i * n + p];
                for (int j : sp.upper[r]) W[(size_t)p * n + j] = -J[(size_t)p * n + j];
            }
            ok = sparse_lu(sp, W.data());
        } else {
            for (int i = 0; i < n * n; ++i) W[i] = -J[i];
            for (int i = 0; i < n; ++i) W[i * n + i] += ig;
            ok = lu_factor(n, W.data(), piv.data());
        }
        st.nlu++;
        if (ok) {
            for (int si = 1; si <= 8; ++si) {
                double* ks = K.data() + si * n;
                if (si == 1) {
                    for (int i = 0; i < n; ++i) rhsv[i] = f0[i] + dt * RD[1] * dT[i];
                } else {
                    for (int i = 0; i < n; ++i) {
                        double acc = u[i];
                        for (int j = 1; j < si; ++j) acc += RA[si][j - 1] * K[j * n + i];
                        us[i] = acc;
                    }
                    f_eval(m, P, s, k, t + RCC[si] * dt, us.data(), rhsv.data());
                    st.nf++;
                    for (int i = 0; i < n; ++i) {
                        double acc = rhsv[i] + dt * RD[si] * dT[i];
                        for (int j = 1; j < si; ++j) acc += RC[si][j - 1] / dt * K[j * n + i];
                        rhsv[i] = acc;
                    }
                }
                if (o.sparse) sparse_solve(sp, W.data(), rhsv.data());
                else lu_solve(n, W.data(), piv.data(), rhsv.data());
                st.nsolve++;
                std::copy(rhsv.begin(), rhsv.end(), ks);
            }
            for (int i = 0; i < n; ++i) u1[i] = us[i] + K[8 * n + i];
        }
        double EEst = INFINITY;
        if (ok) {
            double ss = 0.0;
            bool fin = true;
            for (int i = 0; i < n; ++i) {
                double sc = o.atol + std::max(std::fabs(u[i]), std::fabs(u1[i])) * o.rtol;
                double r = K[8 * n + i] / sc;
                ss += r * r;
                if (!std::isfinite(u1[i])) fin = false;
            }
            EEst = fin ? std::sqrt(ss / n) : INFINITY;
        }
        if (!(EEst <= 1.0)) {
            double q11 = std::isfinite(EEst) ? std::pow(EEst, 0.14) : 5.0;
            dt /= std::min(5.0, q11 / 0.9);
            st.nreject++;
            continue;
        }
        // accept
        double q11 = std::pow(EEst, 0.14);
        double q = (EEst == 0.0) ? 0.1 : std::max(0.1, std::min(5.0, q11 / std::pow(qold, 0.08) / 0.9));
        qold = std::max(EEst, 1e-4);
        double tn = t + dt;
        if (tn >= t1 - 1e-14 * std::fabs(t1)) tn = t1;  // land exactly on the end
        f_eval(m, P, s, k, tn, u1.data(), f1.data());
        st.nf++;
        while (isave < nsave && saveat[isave] <= tn) {
            double th = (saveat[isave] - t) / dt;
            double h00 = (1 + 2 * th) * (1 - th) * (1 - th), h10 = th * (1 - th) * (1 - th);
            double h01 = th * th * (3 - 2 * th), h11 = th * th * (th - 1);
            double* dst = usave + (size_t)isave * n;
            for (int i = 0; i < n; ++i) dst[i] = h00 * u[i] + h10 * dt * f0[i] + h01 * u1[i] + h11 * dt * f1[i];
            ++isave;
        }
        t = tn;
        std::swap(u, u1);
        std::swap(f0, f1);
        st.naccept++;
        for (int i = 0; i < n; ++i)
            if (!std::isfinite(u[i])) { std::copy(u.begin(), u.end(), ufinal); return RET_UNSTABLE; }
        dt = dt / q;
    }
    std::copy(u.begin(), u.end(), ufinal);
    return RET_SUCCESS;
}

}  // namespace ora

// ---------------------------------------------------------------- C API for the tests
using namespace ora;

static Model make_model(const int* md) {
    Model m{md[0], md[1], md[2], md[3], md[4]};
    m.layout();
    return m;
}

extern "C" {

int ora_n_unknowns(const int* md) { return make_model(md).n; }

void ora_layout(const int* md, int* out) {
    Model m = make_model(md);
    int v[9] = {m.iPhi, m.iF, m.iG, m.iN, m.iH, m.iC, m.iB, m.iX, m.n};
    std::memcpy(out, v, sizeof(v));
}

void ora_rhs(const int* md, const double* P, const double* kt, const double* ku, const double* kdu, int nk, double k,
             double tau, const double* u, double* du) {
    Model m = make_model(md);
    Spline s{kt, ku, kdu, nk};
    f_eval(m, P, s, k, tau, u, du);
}

void ora_jacobian(const int* md, const double* P, const double* kt, const double* ku, const double* kdu, int nk,
                  double k, double tau, double* J) {
    Model m = make_model(md);
    Spline s{kt, ku, kdu, nk};
    jacobian(m, P, s, k, tau, J);
}

void ora_dfdt(const int* md, const double* P, const double* kt, const double* ku, const double* kdu, int nk, double k,
              double tau, const double* u, double* out) {
    Model m = make_model(md);
    Spline s{kt, ku, kdu, nk};
    dfdt(m, P, s, k, tau, u, out);
}

void ora_initial_state(const int* md, const double* P, const double* kt, const double* ku, const double* kdu, int nk,
                       double k, double tau, double* u) {
    Model m = make_model(md);
    Spline s{kt, ku, kdu, nk};
    initial_state(m, P, s, k, tau, u);
}

// Background quantities at τ: a, ℋ, κ̇, c_s², ρ_γ, ρ_ν, ρ_c, ρ_b, ρ_h, P_h, κ, ρ_X, w_X, X_e, T_b.
void ora_background(const double* P, const double* kt, const double* ku, const double* kdu, int nk, double tau,
                    double* out) {
    Spline s{kt, ku, kdu, nk};
    Bg<double> b;
    background_at<double>(P, s, tau, b);
    double v[15] = {b.a, b.H, b.kd, b.cs2, b.rg, b.rn, b.rc, b.rb, b.rh, b.ph, b.kappa, b.rx, b.wx, b.Xe, b.Tb};
    std::memcpy(out, v, sizeof(v));
}

double ora_delta_m(const int* md, const double* P, const double* kt, const double* ku, const double* kdu, int nk,
                   double k, double tau, const double* u) {
    Model m = make_model(md);
    Spline s{kt, ku, kdu, nk};
    return delta_m(m, P, s, k, tau, u);
}

double ora_p0(const double* P, double k) { return p0(P, k); }

// (v, v', v'', e^{-κ}) at τ as the sources use them.
void ora_visibility(const double* P, const double* kt, const double* ku, const double* kdu, int nk, double tau,
                    double* out) {
    Spline s{kt, ku, kdu, nk};
    visibility(P, s, tau, out);
}

void ora_sources(const int* md, const double* P, const double* kt, const double* ku, const double* kdu, int nk,
                 double k, double tau, const double* u, double* S) {
    Model m = make_model(md);
    Spline s{kt, ku, kdu, nk};
    sources(m, P, s, k, tau, u, S);
}

// Solve n_modes k-modes with `nthreads` worker threads (dynamic scheduling, one task per mode as
// the reference's @spawn per k, src/solve.jl:566).  stats: 6 int64 per mode.
// With src non-null, also the CMB sources (S_T, S_E, S_ψ) at every saveat time from the saved
// states (src[mode][save][3]), inside the same per-mode task.
void ora_solve_sources(const int* md, const double* P, const double* kt, const double* ku, const double* kdu,
                       int nk, const double* ks, const double* tstart, int n_modes, double rtol, double atol,
                       long maxiters, const double* saveat, int nsave, double* usave, double* ufinal, int* retcode,
                       int64_t* stats, int nthreads, double* src);

void ora_solve(const int* md, const double* P, const double* kt, const double* ku, const double* kdu, int nk,
               const double* ks, const double* tstart, int n_modes, double rtol, double atol, long maxiters,
               const double* saveat, int nsave, double* usave, double* ufinal, int* retcode, int64_t* stats,
               int nthreads) {
    ora_solve_sources(md, P, kt, ku, kdu, nk, ks, tstart, n_modes, rtol, atol, maxiters, saveat, nsave, usave,
                      ufinal, retcode, stats, nthreads, nullptr);
}

static void solve_batch(const int* md, const double* P, const double* kt, const double* ku, const double* kdu,
                        int nk, const double* ks, const double* tstart, int n_modes, double rtol, double atol,
                        long maxiters, const double* saveat, int nsave, double* usave, double* ufinal, int* retcode,
                        int64_t* stats, int nthreads, double* src, bool sparse);

void ora_solve_sources(const int* md, const double* P, const double* kt, const double* ku, const double* kdu,
                       int nk, const double* ks, const double* tstart, int n_modes, double rtol, double atol,
                       long maxiters, const double* saveat, int nsave, double* usave, double* ufinal, int* retcode,
                       int64_t* stats, int nthreads, double* src) {
    solve_batch(md, P, kt, ku, kdu, nk, ks, tstart, n_modes, rtol, atol, maxiters, saveat, nsave, usave, ufinal,
                retcode, stats, nthreads, src, false);
}

// The same batch with the sparse path (colored J, static-pattern LU): the CPU baseline's solver.
void ora_solve_sparse(const int* md, const double* P, const double* kt, const double* ku, const double* kdu, int nk,
                      const double* ks, const double* tstart, int n_modes, double rtol, double atol, long maxiters,
                      const double* saveat, int nsave, double* usave, double* ufinal, int* retcode, int64_t* stats,
                      int nthreads, double* src) {
    solve_batch(md, P, kt, ku, kdu, nk, ks, tstart, n_modes, rtol, atol, maxiters, saveat, nsave, usave, ufinal,
                retcode, stats, nthreads, src, true);
}
}  // extern "C"

static void solve_batch(const int* md, const double* P, const double* kt, const double* ku, const double* kdu,
                        int nk, const double* ks, const double* tstart, int n_modes, double rtol, double atol,
                        long maxiters, const double* saveat, int nsave, double* usave, double* ufinal, int* retcode,
                        int64_t* stats, int nthreads, double* src, bool sparse) {
    Model m = make_model(md);
    Spline s{kt, ku, kdu, nk};
    Opts o{rtol, atol, maxiters};
    o.sparse = sparse;
    double t1 = kt[nk - 1];
    std::atomic<int> next(0);
    auto worker = [&]() {
        for (;;) {
            int i = next.fetch_add(1);
            if (i >= n_modes) break;
            Stats st;
            retcode[i] = solve_mode(m, P, s, ks[i], tstart[i], t1, o, saveat, nsave,
                                    usave ? usave + (size_t)i * nsave * m.n : nullptr, ufinal + (size_t)i * m.n, st);
            int64_t v[6] = {st.naccept, st.nreject, st.nf, st.njac, st.nlu, st.nsolve};
            std::memcpy(stats + 6 * (size_t)i, v, sizeof(v));
            if (src && usave)
                for (int j = 0; j < nsave; ++j)
                    sources(m, P, s, ks[i], saveat[j], usave + ((size_t)i * nsave + j) * m.n,
                            src + ((size_t)i * nsave + j) * 3);
        }
    };
    if (nthreads <= 1) {
        worker();
    } else {
        std::vector<std::thread> th;
        for (int t = 0; t < nthreads; ++t) th.emplace_back(worker);
        for (auto& x : th) x.join();
    }
}
