// Batched Rodas5P perturbation solver for MI355X (gfx950).
//
// One k-mode per lane.  A persistent grid pulls (parameter set, k) instances from a device work
// queue ordered by estimated cost (longest first), so the 10²–10⁵ spread in step counts between
// modes is balanced by lanes fetching new work as soon as their mode finishes.  Each lane runs
// the full Rodas5P step (src/solve.jl:339 → OrdinaryDiffEqRosenbrock) with the model functions
// emitted for this model (RHS, W = I/(γh) - J on the static sparse LU pattern, LU, solves,
// ∂f/∂τ, ICs, Δ_m), spliced in at the @@SBZ_MODEL@@ marker by symboltz.emitter.
//
// Memory layout: every per-mode vector lives in a workspace array indexed [slot][lane], so a
// wavefront's 64 lanes touch 64 consecutive doubles (fully coalesced 512 B); background spline
// tables are read through the cache (one table per parameter set, shared by all its modes).
#ifndef __HIPCC_RTC__  // hiprtc provides the HIP device headers itself
#include <hip/hip_runtime.h>
#endif

#ifndef SBZ_DEV
#define SBZ_DEV __device__ __forceinline__
#endif
#include "../../include/sbz.h"
#include "sbz_dual.h"
#include "sbz_thermo.h"
#include "sbz_rodas5p.h"

// @@SBZ_MODEL@@

#ifndef SBZ_N
#error "model code missing"
#endif

// Views into the workspace.  Each wavefront owns a contiguous block laid out [slot][64 lanes]:
// the lane stride is the compile-time constant 64, so every slot is a distinct constant offset the
// compiler can disambiguate (no false store→load aliasing), and a wavefront's access to one slot
// is 64 consecutive doubles (512 B, fully coalesced).
#define SBZ_WAVE 64
#define SBZ_NB 25
struct SV {
    double* p;
    SBZ_DEV double& operator[](int i) const { return p[i * SBZ_WAVE]; }
};
struct CSV {
    const double* p;
    int s;
    SBZ_DEV double operator[](int i) const { return p[(size_t)i * s]; }
};

struct SbzSolveArgs {
    const double* params;      // [n_sets][SBZ_NPAR]
    const double* knots;       // per set: t[nk], u[nk][5], du[nk][5] at knot_off[set]
    const long long* knot_off;
    const int* knot_n;
    const int* inst_set;       // [n_inst]
    const double* inst_k;      // [n_inst]
    const double* inst_t0;     // [n_inst] start time (already clamped)
    const int* order;          // [n_inst] processing order
    int n_inst;
    double rtol, atol;
    long long maxiters;
    const double* saveat;      // [n_save] shared output times (ascending) or null
    int n_save;
    double* out_u;             // [n_inst][SBZ_N] final state
    double* out_save;          // [n_inst][n_save][SBZ_N] or null
    double* out_src;           // [n_inst][n_save][3] CMB sources (S_T, S_E, S_ψ) or null
    int* retcode;              // [n_inst]
    long long* stats;          // [n_inst][6]: accepted, rejected, f evals, J evals, LU, solves
    unsigned int* queue;       // work counter (zeroed by the host before launch)
    double* ws;                // workspace: per wavefront [SBZ_NW][64]
    int S;                     // number of lanes (multiple of 64)
};

// workspace slots (in units of SBZ_N doubles, then the LU)
#define WS_U0 0
#define WS_U1 1
#define WS_F0 2
#define WS_F1 3
#define WS_DT 4
#define WS_US 5
#define WS_K1 6
#define SBZ_NW (14 * SBZ_N + SBZ_NNZ_LU)
#define WS_W (14 * SBZ_N)

// Background quantities B[] for the perturbation equations at time τ (see symboltz.emitter
// B_NAMES): splined unknowns a, X_H+, X_He+, ΔT, _κ through the cubic Hermite spline, then the
// observed background recomputed from them (sbz_thermo.h).
template <class T>
SBZ_DEV void sbz_pt_background(const double* __restrict__ P, const double* __restrict__ kt, int nk, int& idx,
                               T tau, T* B) {
    const double* ku = kt + nk;
    const double* kdu = kt + 6 * (size_t)nk;
    double tv = dval(tau);
    while (idx > 0 && tv < kt[idx]) --idx;
    while (idx < nk - 2 && tv >= kt[idx + 1]) ++idx;
    double t0 = kt[idx], h = kt[idx + 1] - t0;
    T x = (tau - t0) / h;
    T x2 = x * x, x3 = x2 * x;
    T h00 = 2.0 * x3 - 3.0 * x2 + 1.0, h10 = x3 - 2.0 * x2 + x, h01 = -2.0 * x3 + 3.0 * x2, h11 = x3 - x2;
    T s[5];
#pragma unroll
    for (int j = 0; j < 5; ++j)
        s[j] = h00 * ku[idx * 5 + j] + h10 * (h * kdu[idx * 5 + j]) + h01 * ku[(idx + 1) * 5 + j] +
               h11 * (h * kdu[(idx + 1) * 5 + j]);
    SbzDens<T> d;
    sbz_densities<T>(P, s[0], d);
    SbzThermo<T> th;
    sbz_thermo<T>(P, s[0], s[1], s[2], s[3], d, th);
    B[0] = tau;
    B[1] = s[0];
    B[2] = d.Hc;
    B[3] = th.kappadot;
    B[4] = th.cs2;
    B[5] = d.rho_g;
    B[6] = d.rho_nu;
    B[7] = d.rho_c;
    B[8] = d.rho_b;
    B[9] = d.rho_h;
    B[10] = d.P_h;
    B[11] = d.Irho;
    B[12] = d.IP;
    B[13] = d.rho_de;
    B[14] = d.w_de;
    B[15] = (P[SBZ_P_DE] == 1.0) ? -P[SBZ_P_WA] * s[0] * d.Hc : T(0.0);  // D(w) = -wa a ℋ
#pragma unroll
    for (int i = 0; i < 8; ++i) B[16 + i] = (i < SBZ_NX) ? d.E[i] : T(1.0);
    B[24] = s[4] - P[SBZ_P_KAPPA0];  // κ = _κ - κ0                              baryons.jl:177
}

SBZ_DEV void sbz_bg_plain(const double* P, const double* kt, int nk, int& idx, double tau, double* B) {
    sbz_pt_background<double>(P, kt, nk, idx, tau, B);
}

SBZ_DEV void sbz_bg_dual(const double* P, const double* kt, int nk, int& idx, double tau, double* B, double* dB) {
    Dual Bd[SBZ_NB];
    sbz_pt_background<Dual>(P, kt, nk, idx, Dual(tau, 1.0), Bd);
#pragma unroll
    for (int i = 0; i < SBZ_NB; ++i) {
        B[i] = Bd[i].v;
        dB[i] = Bd[i].d;
    }
}

// Background with first and second τ-derivatives; returns d²κ̇/dτ² (for v̈ in the sources).
SBZ_DEV double sbz_bg_tay2(const double* P, const double* kt, int nk, int& idx, double tau, double* B, double* dB) {
    Tay2 Bt[SBZ_NB];
    sbz_pt_background<Tay2>(P, kt, nk, idx, Tay2(tau, 1.0, 0.0), Bt);
#pragma unroll
    for (int i = 0; i < SBZ_NB; ++i) {
        B[i] = Bt[i].v;
        dB[i] = Bt[i].d;
    }
    return Bt[3].dd;
}

// CMB sources (S_T, S_E, S_ψ) at time ts for the state in slot y; f and g slots are scratch.
template <class SVT>
SBZ_DEV void sbz_sources_at(const double* P, const double* kt, int nk, int& idx, const double* Cs, double k,
                            double ts, SVT y, SVT fy, SVT gy, SVT tmp, double* S) {
    double B[SBZ_NB], dB[SBZ_NB];
    const double ddkd = sbz_bg_tay2(P, kt, nk, idx, ts, B, dB);
    sbz_f(B, Cs, k, y, fy);        // u'
    sbz_f(B, Cs, k, fy, gy);       // J u' (the system is linear)
    sbz_dfdt(B, dB, Cs, k, y, tmp);
    for (int i = 0; i < SBZ_N; ++i) gy[i] = gy[i] + tmp[i];  // u'' = ∂f/∂τ + J u'
    sbz_src(B, dB, ddkd, Cs, k, P[SBZ_P_TAU0], P[SBZ_P_TAUREC], y, fy, gy, S);
}

SBZ_DEV int sbz_find_knot(const double* kt, int nk, double t) {
    int lo = 0, hi = nk - 1;
    if (t <= kt[0]) return 0;
    if (t >= kt[nk - 1]) return nk - 2;
    while (hi - lo > 1) {
        int mid = (lo + hi) >> 1;
        if (kt[mid] <= t) lo = mid; else hi = mid;
    }
    return lo;
}

SBZ_DEV void sbz_consts(const double* P, double* Cs) {
#pragma unroll
    for (int i = 0; i < 8; ++i) {
        double x = (i < SBZ_NX) ? P[SBZ_P_X + i] : 1.0;
        Cs[i] = x;
        Cs[8 + i] = (i < SBZ_NX) ? P[SBZ_P_W + i] : 0.0;
        Cs[16 + i] = -x / (1.0 + exp(-x));  // d ln f0 / d ln x     neutrinos.jl:70
    }
    Cs[24] = P[SBZ_P_CS2X];
}

// Rodas5P tables in constant memory: the stage index is uniform across a wavefront, so these
// are scalar loads.
__constant__ double R5A[9][8] = {
    {0}, {0}, {SBZ_R5_A21}, {SBZ_R5_A31, SBZ_R5_A32}, {SBZ_R5_A41, SBZ_R5_A42, SBZ_R5_A43},
    {SBZ_R5_A51, SBZ_R5_A52, SBZ_R5_A53, SBZ_R5_A54},
    {SBZ_R5_A61, SBZ_R5_A62, SBZ_R5_A63, SBZ_R5_A64, SBZ_R5_A65},
    {SBZ_R5_A61, SBZ_R5_A62, SBZ_R5_A63, SBZ_R5_A64, SBZ_R5_A65, 1.0},
    {SBZ_R5_A61, SBZ_R5_A62, SBZ_R5_A63, SBZ_R5_A64, SBZ_R5_A65, 1.0, 1.0}};
__constant__ double R5C[9][8] = {
    {0}, {0}, {SBZ_R5_C21}, {SBZ_R5_C31, SBZ_R5_C32}, {SBZ_R5_C41, SBZ_R5_C42, SBZ_R5_C43},
    {SBZ_R5_C51, SBZ_R5_C52, SBZ_R5_C53, SBZ_R5_C54},
    {SBZ_R5_C61, SBZ_R5_C62, SBZ_R5_C63, SBZ_R5_C64, SBZ_R5_C65},
    {SBZ_R5_C71, SBZ_R5_C72, SBZ_R5_C73, SBZ_R5_C74, SBZ_R5_C75, SBZ_R5_C76},
    {SBZ_R5_C81, SBZ_R5_C82, SBZ_R5_C83, SBZ_R5_C84, SBZ_R5_C85, SBZ_R5_C86, SBZ_R5_C87}};
__constant__ double R5c[9] = {0, 0, SBZ_R5_C2, SBZ_R5_C3, SBZ_R5_C4, SBZ_R5_C5, 1.0, 1.0, 1.0};
__constant__ double R5d[9] = {0, SBZ_R5_D1, SBZ_R5_D2, SBZ_R5_D3, SBZ_R5_D4, SBZ_R5_D5, 0, 0, 0};

extern "C" __global__ void __launch_bounds__(64) sbz_rodas5p_kernel(SbzSolveArgs a) {
    const int lane = blockIdx.x * blockDim.x + threadIdx.x;
    if (lane >= a.S) return;
    // this wavefront's workspace block
    double* wsl = a.ws + (size_t)(lane / SBZ_WAVE) * SBZ_NW * SBZ_WAVE + (lane % SBZ_WAVE);
    auto vec = [&](int slot) { return SV{wsl + (size_t)slot * SBZ_N * SBZ_WAVE}; };
    const SV W{wsl + (size_t)WS_W * SBZ_WAVE};


    for (;;) {
        const unsigned int qi = atomicAdd(a.queue, 1u);
        if (qi >= (unsigned int)a.n_inst) break;
        const int inst = a.order ? a.order[qi] : (int)qi;
        const int set = a.inst_set[inst];
        const double* P = a.params + (size_t)set * SBZ_NPAR;
        const double* kt = a.knots + a.knot_off[set];
        const int nk = a.knot_n[set];
        const double k = a.inst_k[inst];
        const double t1 = kt[nk - 1];
        double t = a.inst_t0[inst];
        double Cs[25];
        sbz_consts(P, Cs);
        double B[SBZ_NB], dB[SBZ_NB];
        int kidx = sbz_find_knot(kt, nk, t);
        long long st_acc = 0, st_rej = 0, st_f = 0, st_j = 0;
        int ret = SBZ_RET_SUCCESS;
        int isave = 0;
        double* osave = a.out_save ? a.out_save + (size_t)inst * a.n_save * SBZ_N : nullptr;
        int su = WS_U0, su1 = WS_U1, sf0 = WS_F0, sf1 = WS_F1;

        // initial conditions
        sbz_bg_plain(P, kt, nk, kidx, t, B);
        {
            SV u = vec(su);
            sbz_ic(B, Cs, k, u);
        }
        bool finite0 = (k > 0.0) && (k == k);
        for (int i = 0; i < SBZ_N; ++i) {
            double v = vec(su)[i];
            if (!(v == v) || fabs(v) > 1e300) finite0 = false;
        }
        if (!finite0) ret = SBZ_RET_INITIALFAILURE;
        double* osrc = a.out_src ? a.out_src + (size_t)inst * a.n_save * 3 : nullptr;
        while ((osave || osrc) && isave < a.n_save && a.saveat[isave] <= t) {
            if (osave)
                for (int i = 0; i < SBZ_N; ++i) osave[(size_t)isave * SBZ_N + i] = vec(su)[i];
            if (osrc && ret == SBZ_RET_SUCCESS) {
                double S3[3];
                sbz_sources_at(P, kt, nk, kidx, Cs, k, a.saveat[isave], vec(su), vec(WS_K1), vec(WS_K1 + 1),
                               vec(WS_K1 + 2), S3);
                for (int c = 0; c < 3; ++c) osrc[(size_t)isave * 3 + c] = S3[c];
            }
            ++isave;
        }
        double dt = 0.0, qold = SBZ_CTL_QOLDINIT;
        if (ret == SBZ_RET_SUCCESS && t1 > t) {
            // Hairer–Wanner initial step (OrdinaryDiffEqCore ode_determine_initdt)
            SV u = vec(su), f0 = vec(sf0), u1 = vec(su1), f1 = vec(sf1);
            sbz_f(B, Cs, k, u, f0);
            double d0 = 0.0, d1 = 0.0;
            for (int i = 0; i < SBZ_N; ++i) {
                double sk = a.atol + fabs(u[i]) * a.rtol;
                d0 += (u[i] / sk) * (u[i] / sk);
                d1 += (f0[i] / sk) * (f0[i] / sk);
            }
            d0 = sqrt(d0 / SBZ_N);
            d1 = sqrt(d1 / SBZ_N);
            double dt0 = (d0 < 1e-5 || d1 < 1e-5) ? 1e-6 : 0.01 * d0 / d1;
            dt0 = fmin(dt0, t1 - t);
            for (int i = 0; i < SBZ_N; ++i) u1[i] = u[i] + dt0 * f0[i];
            sbz_bg_plain(P, kt, nk, kidx, t + dt0, B);
            sbz_f(B, Cs, k, u1, f1);
            double d2 = 0.0;
            for (int i = 0; i < SBZ_N; ++i) {
                double sk = a.atol + fabs(u[i]) * a.rtol;
                double r = (f1[i] - f0[i]) / sk;
                d2 += r * r;
            }
            d2 = sqrt(d2 / SBZ_N) / dt0;
            double mx = fmax(d1, d2);
            double dt1 = (mx <= 1e-15) ? fmax(1e-6, dt0 * 1e-3) : pow(0.01 / mx, 1.0 / 6.0);
            dt = fmin(fmin(100.0 * dt0, dt1), t1 - t);
            st_f += 2;
        }
        long long iters = 0;
        while (ret == SBZ_RET_SUCCESS && t < t1) {
            if (++iters > a.maxiters) { ret = SBZ_RET_MAXITERS; break; }
            if (t + dt >= t1 - 1e-12 * fabs(t1)) dt = t1 - t;
            else if (dt < 1e-14 * fabs(t)) { ret = SBZ_RET_DTLESSTHANMIN; break; }
            SV u = vec(su), u1 = vec(su1), f0 = vec(sf0), f1 = vec(sf1), dT = vec(WS_DT), us = vec(WS_US);
            // J, ∂f/∂τ and W at (t, u)
            sbz_bg_dual(P, kt, nk, kidx, t, B, dB);
            const double ig = 1.0 / (SBZ_R5_GAMMA * dt);
            sbz_w(B, Cs, k, ig, W);
            bool ok = sbz_lu(W);
            sbz_dfdt(B, dB, Cs, k, u, dT);
            st_j++;
            double EEst = __builtin_inf();
            if (ok) {
                double tprev = -1.0;
                for (int s = 1; s <= 8; ++s) {
                    SV ks = vec(WS_K1 + s - 1);
                    if (s == 1) {
                        for (int i = 0; i < SBZ_N; ++i) ks[i] = f0[i] + dt * SBZ_R5_D1 * dT[i];
                    } else {
                        for (int i = 0; i < SBZ_N; ++i) {
                            double acc = u[i];
                            for (int j = 1; j < s; ++j) acc += R5A[s][j - 1] * vec(WS_K1 + j - 1)[i];
                            us[i] = acc;
                        }
                        const double ts = t + R5c[s] * dt;
                        if (ts != tprev) {
                            sbz_bg_plain(P, kt, nk, kidx, ts, B);
                            tprev = ts;
                        }
                        sbz_f(B, Cs, k, us, ks);
                        st_f++;
                        const double dd = dt * R5d[s];
                        for (int i = 0; i < SBZ_N; ++i) {
                            double acc = ks[i] + dd * dT[i];
                            for (int j = 1; j < s; ++j) acc += (R5C[s][j - 1] / dt) * vec(WS_K1 + j - 1)[i];
                            ks[i] = acc;
                        }
                    }
                    sbz_solve(W, ks);
                }
                SV k8 = vec(WS_K1 + 7);
                double ss = 0.0;
                bool fin = true;
                for (int i = 0; i < SBZ_N; ++i) {
                    double un = us[i] + k8[i];
                    u1[i] = un;
                    double sc = a.atol + fmax(fabs(u[i]), fabs(un)) * a.rtol;
                    double r = k8[i] / sc;
                    ss += r * r;
                    if (!(un == un) || fabs(un) > 1e300) fin = false;
                }
                EEst = fin ? sqrt(ss / SBZ_N) : __builtin_inf();
            }
            if (!(EEst <= 1.0)) {
                double q11 = (EEst < __builtin_inf()) ? pow(EEst, SBZ_CTL_BETA1) : 1.0 / SBZ_CTL_QMIN;
                dt /= fmin(1.0 / SBZ_CTL_QMIN, q11 / SBZ_CTL_GAMMA);
                st_rej++;
                continue;
            }
            double q11 = pow(EEst, SBZ_CTL_BETA1);
            double q = (EEst == 0.0) ? 1.0 / SBZ_CTL_QMAX
                                     : fmax(1.0 / SBZ_CTL_QMAX,
                                            fmin(1.0 / SBZ_CTL_QMIN, q11 / pow(qold, SBZ_CTL_BETA2) / SBZ_CTL_GAMMA));
            qold = fmax(EEst, SBZ_CTL_QOLDINIT);
            double tn = t + dt;
            if (tn >= t1 - 1e-14 * fabs(t1)) tn = t1;
            // f at the new point: next step's first stage and the Hermite dense output
            sbz_bg_plain(P, kt, nk, kidx, tn, B);
            sbz_f(B, Cs, k, u1, f1);
            st_f++;
            while ((osave || osrc) && isave < a.n_save && a.saveat[isave] <= tn) {
                // cubic Hermite dense output over the accepted step
                double th = (a.saveat[isave] - t) / dt;
                double h00 = (1 + 2 * th) * (1 - th) * (1 - th), h10 = th * (1 - th) * (1 - th);
                double h01 = th * th * (3 - 2 * th), h11 = th * th * (th - 1);
                SV ys = vec(WS_K1);
                for (int i = 0; i < SBZ_N; ++i) ys[i] = h00 * u[i] + h10 * dt * f0[i] + h01 * u1[i] + h11 * dt * f1[i];
                if (osave)
                    for (int i = 0; i < SBZ_N; ++i) osave[(size_t)isave * SBZ_N + i] = ys[i];
                if (osrc) {
                    double S3[3];
                    sbz_sources_at(P, kt, nk, kidx, Cs, k, a.saveat[isave], ys, vec(WS_K1 + 1), vec(WS_K1 + 2),
                                   vec(WS_K1 + 3), S3);
                    for (int c = 0; c < 3; ++c) osrc[(size_t)isave * 3 + c] = S3[c];
                }
                ++isave;
            }
            t = tn;
            int tmp = su; su = su1; su1 = tmp;
            tmp = sf0; sf0 = sf1; sf1 = tmp;
            st_acc++;
            dt = dt / q;
        }
        {
            SV u = vec(su);
            double* o = a.out_u + (size_t)inst * SBZ_N;
            for (int i = 0; i < SBZ_N; ++i) o[i] = u[i];
        }
        a.retcode[inst] = ret;
        long long* st = a.stats + (size_t)inst * 6;
        st[0] = st_acc;
        st[1] = st_rej;
        st[2] = st_f;
        st[3] = st_j;
        st[4] = st_j;
        st[5] = 8 * st_j;
    }
}

// P(k, τ0) = P0(k) Δ_m(τ0)² for each instance from its final state (fourier.jl:14-52,90-97).
extern "C" __global__ void sbz_pk_kernel(const double* __restrict__ params, const double* __restrict__ knots,
                                         const long long* __restrict__ knot_off, const int* __restrict__ knot_n,
                                         const int* __restrict__ inst_set, const double* __restrict__ inst_k,
                                         const double* __restrict__ u_final, int n_inst, double* __restrict__ pk) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n_inst) return;
    int set = inst_set[i];
    const double* P = params + (size_t)set * SBZ_NPAR;
    const double* kt = knots + knot_off[set];
    int nk = knot_n[set];
    double k = inst_k[i];
    double Cs[25], B[SBZ_NB];
    sbz_consts(P, Cs);
    int idx = nk - 2;
    sbz_bg_plain(P, kt, nk, idx, kt[nk - 1], B);
    CSV u{u_final + (size_t)i * SBZ_N, 1};
    double dm = sbz_dm(B, Cs, k, u);
    double p0 = 2.0 * sbzc::PI * sbzc::PI / (k * k * k) * P[SBZ_P_AS] * pow(k / P[SBZ_P_KPIVOT], P[SBZ_P_NS] - 1.0);
    pk[i] = p0 * dm * dm;
}

#ifdef SBZ_HOST_PROBE
// Host entry points into the emitted model functions, compiled only into the test probe library
// (SBZ_DEV = __host__ __device__): they let the CPU test-suite compare the emitted RHS, W, LU,
// solves, ∂f/∂τ, initial conditions and Δ_m with the oracle without a GPU.
struct HV {
    double* p;
    double& operator[](int i) const { return p[i]; }
};
extern "C" int sbz_probe_n(void) { return SBZ_N; }
extern "C" int sbz_probe_nnz(void) { return SBZ_NNZ_LU; }
extern "C" void sbz_probe_bg(const double* P, const double* kt, int nk, double tau, double* B, double* dB) {
    int idx = sbz_find_knot(kt, nk, tau);
    Dual Bd[SBZ_NB];
    sbz_pt_background<Dual>(P, kt, nk, idx, Dual(tau, 1.0), Bd);
    for (int i = 0; i < SBZ_NB; ++i) { B[i] = Bd[i].v; dB[i] = Bd[i].d; }
}
extern "C" void sbz_probe_src(const double* P, const double* kt, int nk, double k, double tau, const double* u,
                              double* S, double* work) {
    double Cs[25]; sbz_consts(P, Cs);
    int idx = sbz_find_knot(kt, nk, tau);
    for (int i = 0; i < SBZ_N; ++i) work[i] = u[i];
    sbz_sources_at(P, kt, nk, idx, Cs, k, tau, HV{work}, HV{work + SBZ_N}, HV{work + 2 * SBZ_N},
                   HV{work + 3 * SBZ_N}, S);
}
extern "C" void sbz_probe_f(const double* B, const double* P, double k, const double* u, double* du) {
    double Cs[25]; sbz_consts(P, Cs);
    sbz_f(B, Cs, k, HV{(double*)u}, HV{du});
}
extern "C" void sbz_probe_w(const double* B, const double* P, double k, double ig, double* W) {
    double Cs[25]; sbz_consts(P, Cs);
    sbz_w(B, Cs, k, ig, HV{W});
}
extern "C" int sbz_probe_lu(double* W) { return sbz_lu(HV{W}) ? 1 : 0; }
extern "C" void sbz_probe_solve(const double* W, double* b) { sbz_solve(HV{(double*)W}, HV{b}); }
extern "C" void sbz_probe_dfdt(const double* B, const double* dB, const double* P, double k, const double* u, double* out) {
    double Cs[25]; sbz_consts(P, Cs);
    sbz_dfdt(B, dB, Cs, k, HV{(double*)u}, HV{out});
}
extern "C" void sbz_probe_ic(const double* B, const double* P, double k, double* u) {
    double Cs[25]; sbz_consts(P, Cs);
    sbz_ic(B, Cs, k, HV{u});
}
extern "C" double sbz_probe_dm(const double* B, const double* P, double k, const double* u) {
    double Cs[25]; sbz_consts(P, Cs);
    return sbz_dm(B, Cs, k, HV{(double*)u});
}
#endif
