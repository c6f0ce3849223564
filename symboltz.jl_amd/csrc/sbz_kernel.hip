// Batched Rodas5P perturbation solver for MI355X (gfx950).
//
// A persistent grid pulls (parameter set, k) instances from device work queues (one per XCD,
// each ordered by estimated cost, longest first), so the 10²–10⁵ spread in step counts between
// modes is balanced by wavefronts fetching new work as soon as their mode finishes.  Every kernel
// runs the full Rodas5P step (src/solve.jl:339 → OrdinaryDiffEqRosenbrock) on the model functions
// emitted for this model and spliced in at the @@SBZ_MODEL@@ marker by symboltz.emitter:
//   sbz_wave_kernel, sbz_wave_kernel_o2   one mode per 64-lane wavefront (one / two per SIMD)
//   sbz_wave4_kernel                      four modes per wavefront, 16 lanes each
//   sbz_wave2h_kernel                     two modes per wavefront, 32 lanes each
//   sbz_wave_tan_kernel                   one mode and its parameter tangent per wavefront
// plus the source background table, P(k) and output-fill kernels.  Background spline tables (one
// per parameter set, shared by all its modes) are read through the cache.
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

// Broadcast a double from a (wave-uniform) lane.
SBZ_DEV double sbz_readlane(double v, int l) {
    const long long x = __double_as_longlong(v);
    const int lo = __builtin_amdgcn_readlane((int)(x & 0xffffffffll), l);
    const int hi = __builtin_amdgcn_readlane((int)(x >> 32), l);
    return __longlong_as_double((long long)(((unsigned long long)(unsigned int)hi << 32) | (unsigned int)lo));
}

// DPP lane permutation of a double within rows of 16 lanes (all 64 lanes must be active).
template <int CTRL>
SBZ_DEV double sbz_dpp(double v) {
    const long long x = __double_as_longlong(v);
    const int lo = __builtin_amdgcn_mov_dpp((int)(x & 0xffffffffll), CTRL, 0xF, 0xF, false);
    const int hi = __builtin_amdgcn_mov_dpp((int)(x >> 32), CTRL, 0xF, 0xF, false);
    return __longlong_as_double((long long)(((unsigned long long)(unsigned int)hi << 32) | (unsigned int)lo));
}

// Broadcast classes of the emitted border LU and triangular solves (BC::at<J>(v): lane J's v):
// over the whole wavefront (one mode per wavefront) or within each row of 16 lanes (row_newbcast,
// four modes per wavefront, one per row).
struct SbzBcWave {
    template <int J>
    static SBZ_DEV double at(double v) { return sbz_readlane(v, J); }
};
struct SbzBcRow16 {
    template <int J>
    static SBZ_DEV double at(double v) { return sbz_dpp<0x150 + J>(v); }
};
// Within each half of 32 lanes (two modes per wavefront): lane J and lane 32 + J read out, each
// half keeping its own.
struct SbzBcHalf32 {
    template <int J>
    static SBZ_DEV double at(double v) {
        const double lo = sbz_readlane(v, J), hi = sbz_readlane(v, 32 + J);
        return (threadIdx.x & 32) ? hi : lo;
    }
};

// @@SBZ_MODEL@@

#ifndef SBZ_N
#error "model code missing"
#endif

#define SBZ_WAVE 64
#define SBZ_OWN ((SBZ_N + 63) / 64)   // state components per lane in the one-mode-per-wavefront kernels
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
    const double* saveat;      // output times (ascending): [n_save] shared, or [n_sets][n_save] per set
    int n_save;
    int saveat_stride;         // 0 (shared) or n_save (one row per parameter set)
    double* out_u;             // [n_inst][SBZ_N] final state
    double* out_save;          // [n_inst][n_save][SBZ_N] or null
    double* out_src;           // [n_inst][n_save][3] CMB sources (S_T, S_E, S_ψ) or null
    int* retcode;              // [n_inst]
    long long* stats;          // [n_inst][6]: accepted, rejected, f evals, J evals, LU, solves
    unsigned int* queue;       // work counter (zeroed by the host before launch)
    int S;                     // number of resident wavefronts (the persistent grid)
    const double* srcbg;       // [n_sets][n_save][SBZ_SRCBG_STRIDE] background at the save times
                               // (sbz_srcbg_kernel), or null: the wave kernel then evaluates it per mode
    const int* xcd_off;        // [n_xcd + 1] segments of `order`, one work queue per XCD; null: one queue
    int n_xcd;                 // number of queues (queue[0..n_xcd) counters)
    // tangent solves (sbz_wave_tan_kernel): per instance the parameter sets of log θ + h and
    // log θ - h and 1/(2h); outputs s = ∂u/∂log θ at τ0 and f(τ0, u(τ0))
    const int* inst_setp;
    const int* inst_setm;
    const double* inst_tscale;
    double* out_s;             // [n_inst][SBZ_N]
    double* out_fend;          // [n_inst][SBZ_N]
};

// Next position in `order` for a wavefront on XCD x (blockIdx mod n_xcd: workgroups are dispatched
// round-robin over the XCDs), or -1 when all work is taken.  With per-XCD queues each XCD first
// drains its own segment (the instances of the parameter sets assigned to it, so their knot tables
// stay in that XCD's L2), then takes work from the other segments.
SBZ_DEV int sbz_pull(const SbzSolveArgs& a, int x) {
    if (!a.xcd_off) {
        const unsigned int qi = atomicAdd(a.queue, 1u);
        return qi < (unsigned int)a.n_inst ? (int)qi : -1;
    }
    for (int d = 0; d < a.n_xcd; ++d) {
        const int q = (x + d) % a.n_xcd;
        const int lo = a.xcd_off[q], cnt = a.xcd_off[q + 1] - lo;
        if (cnt <= 0) continue;
        const unsigned int qi = atomicAdd(a.queue + q, 1u);
        if (qi < (unsigned int)cnt) return lo + (int)qi;
    }
    return -1;
}

// Row of the per-set source background table: B[0..NB), dB/dτ[0..NB), d²κ̇/dτ² at 2 NB.
#define SBZ_SRCBG_STRIDE 80

// Background quantities B[] for the perturbation equations at time τ (see symboltz.emitter
// B_NAMES): splined unknowns a, X_H+, X_He+, ΔT, _κ through the cubic Hermite spline, then the
// observed background recomputed from them (sbz_thermo.h).
template <class T>
SBZ_DEV void sbz_pt_background(const double* __restrict__ P, const double* __restrict__ kt, int nk, int& idx,
                               T tau, T* B) {
    const double* ku = kt + nk;
    const double* kdu = kt + 6 * (size_t)nk;
    double tv = dval(tau);
    // The interval of the cursor and its spline data are loaded together: τ usually still lies in
    // it, and then one round trip to memory serves both the search and the spline.  Otherwise the
    // cursor walks to τ's interval and the data are loaded again.  (idx stays in [0, nk - 2], and
    // every set has nk >= 2 knots.)
    double ta = kt[idx], tb = kt[idx + 1];
    double ua[5], ub[5], da[5], db[5];
#pragma unroll
    for (int j = 0; j < 5; ++j) {
        ua[j] = ku[idx * 5 + j];
        ub[j] = ku[(idx + 1) * 5 + j];
        da[j] = kdu[idx * 5 + j];
        db[j] = kdu[(idx + 1) * 5 + j];
    }
    if ((idx > 0 && tv < ta) || (idx < nk - 2 && tv >= tb)) {
        while (idx > 0 && tv < kt[idx]) --idx;
        while (idx < nk - 2 && tv >= kt[idx + 1]) ++idx;
        ta = kt[idx];
        tb = kt[idx + 1];
#pragma unroll
        for (int j = 0; j < 5; ++j) {
            ua[j] = ku[idx * 5 + j];
            ub[j] = ku[(idx + 1) * 5 + j];
            da[j] = kdu[idx * 5 + j];
            db[j] = kdu[(idx + 1) * 5 + j];
        }
    }
    double t0 = ta, h = tb - t0;
    T x = (tau - t0) / h;
    T x2 = x * x, x3 = x2 * x;
    T h00 = 2.0 * x3 - 3.0 * x2 + 1.0, h10 = x3 - 2.0 * x2 + x, h01 = -2.0 * x3 + 3.0 * x2, h11 = x3 - x2;
    T s[5];
#pragma unroll
    for (int j = 0; j < 5; ++j) s[j] = h00 * ua[j] + h10 * (h * da[j]) + h01 * ub[j] + h11 * (h * db[j]);
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
    // reciprocals, once per evaluation, for the wavefront kernel's division-free expressions
    // (values only: the emitted ∂f/∂τ is differentiated before the reciprocals are substituted,
    // so no derivative of a reciprocal is ever read)
    B[SBZ_B_IHC] = T(1.0 / dval(d.Hc));
    B[SBZ_B_IIRHOP] = T(1.0 / dval(d.Irho + d.IP / 3.0));
    B[SBZ_B_IIRHO] = T(1.0 / dval(d.Irho));
    B[SBZ_B_IRB] = T(1.0 / dval(d.rho_b));
    B[SBZ_B_IWX1] = T((P[SBZ_P_DE] == 1.0) ? 1.0 / (1.0 + dval(d.w_de)) : 0.0);
    B[SBZ_B_ITAU] = T(1.0 / dval(tau));
#pragma unroll
    for (int i = 0; i < 8; ++i) B[SBZ_B_IE0 + i] = T((i < SBZ_NX) ? 1.0 / dval(d.E[i]) : 1.0);
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

// Row g = set * n_save + i of the source background table: the Taylor-2 background at save time i
// of one parameter set.  It depends on the set and τ only, so one evaluation serves every mode of
// the set.  The knot search in sbz_pt_background ends on the same interval from any start, so the
// values are those the wave kernel would compute at the save.
static_assert(2 * SBZ_NB + 1 <= SBZ_SRCBG_STRIDE, "source background row too short");
SBZ_DEV void sbz_srcbg_row(const double* params, const double* knots, const long long* knot_off, const int* knot_n,
                           const double* saveat, int saveat_stride, int n_save, long long g, double* out) {
    const int set = (int)(g / n_save), i = (int)(g % n_save);
    const double* P = params + (size_t)set * SBZ_NPAR;
    const double* kt = knots + knot_off[set];
    const int nk = knot_n[set];
    const double tau = saveat[(size_t)set * saveat_stride + i];
    int kk = sbz_find_knot(kt, nk, tau);
    double B[SBZ_NB], dB[SBZ_NB];
    const double dd = sbz_bg_tay2(P, kt, nk, kk, tau, B, dB);
    double* o = out + (size_t)g * SBZ_SRCBG_STRIDE;
#pragma unroll
    for (int q = 0; q < SBZ_NB; ++q) {
        o[q] = B[q];
        o[SBZ_NB + q] = dB[q];
    }
    o[2 * SBZ_NB] = dd;
}

SBZ_DEV void sbz_consts(const double* P, double* Cs) {
#pragma unroll
    for (int i = 0; i < 8; ++i) {
        double x = (i < SBZ_NX) ? P[SBZ_P_X + i] : 1.0;
        Cs[i] = x;
        Cs[8 + i] = (i < SBZ_NX) ? P[SBZ_P_W + i] : 0.0;
        Cs[16 + i] = -x / (1.0 + exp(-x));  // d ln f0 / d ln x     neutrinos.jl:70
        Cs[SBZ_C_IX0 + i] = 1.0 / x;
    }
    Cs[24] = P[SBZ_P_CS2X];
    Cs[SBZ_C_IK] = 0.0;   // set per mode (sbz_consts_k)
    Cs[SBZ_C_IK2] = 0.0;
}
SBZ_DEV void sbz_consts_k(double* Cs, double k) {
    Cs[SBZ_C_IK] = 1.0 / k;
    Cs[SBZ_C_IK2] = 1.0 / (k * k);
}

// ============================================================================================
// Wavefront-per-mode Rodas5P kernel.
//
// One 64-lane wavefront integrates one k-mode.  The unknowns are split (symboltz.emitter
// wave_structure) into chains — the tridiagonal interiors l = 3..L of every hierarchy, one
// chain per lane, eliminated by register-resident sweeps (top-down: the stable direction, every
// modified pivot stays ≥ 1/(γh)) — and ≤ 64 border unknowns (Φ, l ≤ 2 multipoles, fluids), one per
// lane, whose W block is factored row-parallel with the pivot row broadcast through LDS and
// solved with readlane broadcasts.  A chain couples to the border only through its l = 2
// attachment, so its Schur complement is a single diagonal correction.  Stage vectors live in
// LDS; the background at the six distinct stage times is evaluated by six lanes at once.
// ============================================================================================
__constant__ int sbz_stage_bg[10] = {0, 0, 1, 2, 3, 4, 5, 5, 5, 5};  // stage -> background slot

// The border staging (A, α, β, p, q) lives in the stage vectors K1..K8 when it fits there: they
// are dead while W is assembled at the start of a step attempt, and the staging is dead once the
// rows are loaded (its loads mask every entry the assembly does not write).
#define SBZ_M0_IN_K (SBZ_NBORD * SBZ_WSTRIDE <= 8 * SBZ_N)
template <bool SPLIT>
struct WaveLDSt {
    double v[14][SBZ_N];           // 0,1 U (ping-pong), 2,3 F (ping-pong), 4 DT, 5 US, 6..13 K1..K8
    double B[7][SBZ_NB];           // background at the distinct stage times (0..5) and a spare slot (6)
    double dB[2][SBZ_NB];          // d/dτ of the background at t_n (0) and at the spare time (1)
    double Cs[SBZ_NC];             // per-mode constants (momentum nodes, weights, d ln f0/d ln x, c_s², 1/x, 1/k)
#if !SBZ_M0_IN_K
    double M0[SBZ_NBORD * SBZ_WSTRIDE];   // border staging
#endif
    double LU[(1 + SPLIT) * SBZ_NBORD * SBZ_WSTRIDE];   // factored border block (sbz_wb_store<SPLIT>)
    double X[SBZ_NBORD];           // border solution broadcast
    double C[SBZ_NCHAIN];          // chain → border exchange
    double ddkd;
    int mode;
};
using WaveLDS = WaveLDSt<SBZ_LU_SPLIT>;   // the model's default factor layout
template <bool SPLIT>
SBZ_DEV double* sbz_wave_m0(WaveLDSt<SPLIT>& L) {
#if SBZ_M0_IN_K
    return &L.v[6][0];
#else
    return L.M0;
#endif
}

// select element `i` of an 8-vector with compile-time indices (no scratch for runtime indexing)
SBZ_DEV double sbz_pick8(const double* v, int i) {
    double r = v[0];
#pragma unroll
    for (int j = 1; j < 8; ++j) r = (i == j) ? v[j] : r;
    return r;
}

// Chain rows.  Row l of a chain (3 <= l <= L) reads  du_l = ca u_{l-1} + cd u_l + cb u_{l+1}  with
//   l < L:  ca =  s k l/(2l+1),  cb = -s k (l+1)/(2l+1),  cd = dd
//   l = L:  ca =  s k,           cb = 0,                  cd = dd - (L+1)/τ
// (photons.jl:32-33,58-59, neutrinos.jl:25-26,131-132), where dd = κ̇ (B[3], negative in the
// reference's convention) for photons and 0 otherwise, and s = x_i/E_i for massive-ν bin i
// (1 for massless species).  The loops over l are unrolled, so l/(2l+1) and (l+1)/(2l+1) are
// compile-time constants.  SbzChainRow holds the l-independent per-lane part at one time.
struct WaveCtx {
    int lane, isB, bcomp, isC, ckind, cL, cbin, catt, cattc, bchain, cbase;
};
struct SbzChainRow {
    double sk;   // s k
    double dd;   // diagonal of the rows l < L
    double top;  // diagonal of the row l = L
};
SBZ_DEV SbzChainRow sbz_chain_row(const WaveCtx& c, const double* B, const double* Cs, double k) {
    SbzChainRow r;
    r.sk = (c.ckind == 2) ? k * sbz_pick8(Cs, c.cbin) * sbz_pick8(B + SBZ_B_IE0, c.cbin) : k;
    r.dd = (c.ckind == 0) ? B[3] : 0.0;
    r.top = r.dd - (c.cL + 1.0) * B[SBZ_B_ITAU];
    return r;
}
// the same coefficients differentiated in τ at fixed u
SBZ_DEV SbzChainRow sbz_chain_drow(const WaveCtx& c, const double* B, const double* dB, const double* Cs, double k) {
    SbzChainRow r;
    const double iE = sbz_pick8(B + SBZ_B_IE0, c.cbin), it = B[SBZ_B_ITAU];
    r.sk = (c.ckind == 2) ? -k * sbz_pick8(Cs, c.cbin) * sbz_pick8(dB + 16, c.cbin) * (iE * iE) : 0.0;
    r.dd = (c.ckind == 0) ? dB[3] : 0.0;
    r.top = r.dd + (c.cL + 1.0) * (it * it);
    return r;
}
#define SBZ_RL(e) ((3.0 + (e)) / (2.0 * (3.0 + (e)) + 1.0))    // l/(2l+1) at l = 3 + e
#define SBZ_RL1(e) ((4.0 + (e)) / (2.0 * (3.0 + (e)) + 1.0))   // (l+1)/(2l+1) at l = 3 + e

// Chain kernels, compiled per chain length LEN (chains of one length share a lane group, so
// their loops have compile-time bounds and no per-lane predicates).  Element e holds l = 3 + e.
//   rows:   out_e = ca_e u_{e-1} + cd_e u_e + cb_e u_{e+1},  u_{-1} = attachment (l = 2)
//   factor: d'_e = W[e][e] - W[e][e+1] W[e+1][e] / d'_e+1 from the top (l = L) down, W = I/(γh) - J
//   fwd:    b'_e = b_e - W[e][e+1] b'_{e+1} / d'_{e+1}          (in place)
//   bwd:    x_e  = (b'_e - W[e][e-1] x_{e-1}) / d'_e              (in place, x_{-1} = attachment)
template <int LEN>
SBZ_DEV void sbz_chain_rows(const SbzChainRow& r, double att, const double* ci, double* co) {
    double v[LEN + 2];
    v[0] = att;
#pragma unroll
    for (int e = 0; e < LEN; ++e) v[e + 1] = ci[e];
    v[LEN + 1] = 0.0;
#pragma unroll
    for (int e = 0; e < LEN - 1; ++e) co[e] = r.sk * SBZ_RL(e) * v[e] + r.dd * v[e + 1] - r.sk * SBZ_RL1(e) * v[e + 2];
    co[LEN - 1] = r.sk * v[LEN - 1] + r.top * v[LEN];
}
template <int LEN>
SBZ_DEV double sbz_chain_factor(const SbzChainRow& r, double ig, double* cinv) {
    double ca_next = r.sk;   // -W[e+1][e]; the top row's is s k
    cinv[LEN - 1] = 1.0 / (ig - r.top);
#pragma unroll
    for (int e = LEN - 2; e >= 0; --e) {
        cinv[e] = 1.0 / (ig - r.dd + r.sk * SBZ_RL1(e) * ca_next * cinv[e + 1]);
        ca_next = r.sk * SBZ_RL(e);
    }
    return -ca_next * cinv[0];   // W[0][att] / d'_0: the Schur term of the attachment row
}
template <int LEN>
SBZ_DEV double sbz_chain_fwd(double sk, const double* cinv, double* kc) {
    double b[LEN];
#pragma unroll
    for (int e = 0; e < LEN; ++e) b[e] = kc[e];
#pragma unroll
    for (int e = LEN - 2; e >= 0; --e) b[e] -= sk * SBZ_RL1(e) * cinv[e + 1] * b[e + 1];
#pragma unroll
    for (int e = 0; e < LEN; ++e) kc[e] = b[e];
    return b[0] * cinv[0];
}
template <int LEN>
SBZ_DEV void sbz_chain_bwd(double sk, const double* cinv, double xatt, double* kc) {
    double b[LEN];
#pragma unroll
    for (int e = 0; e < LEN; ++e) b[e] = kc[e];
    double prev = xatt;
#pragma unroll
    for (int e = 0; e < LEN; ++e) {
        prev = (b[e] + ((e + 1 == LEN) ? sk : sk * SBZ_RL(e)) * prev) * cinv[e];
        b[e] = prev;
    }
#pragma unroll
    for (int e = 0; e < LEN; ++e) kc[e] = b[e];
}
#define SBZ_IN_GROUP(lo, hi) (c.lane >= (lo) && c.lane < (hi))

// The chain rows with the chain's inputs already in registers (cv[e] = the member l = 3 + e).
template <int LEN>
SBZ_DEV void sbz_chain_rows_r(const SbzChainRow& r, double att, const double* cv, double* co) {
    double v[LEN + 2];
    v[0] = att;
#pragma unroll
    for (int e = 0; e < LEN; ++e) v[e + 1] = cv[e];
    v[LEN + 1] = 0.0;
#pragma unroll
    for (int e = 0; e < LEN - 1; ++e) co[e] = r.sk * SBZ_RL(e) * v[e] + r.dd * v[e + 1] - r.sk * SBZ_RL1(e) * v[e + 2];
    co[LEN - 1] = r.sk * v[LEN - 1] + r.top * v[LEN];
}

// out = J(τ) in, i.e. f(τ, in) (f is linear in u); with dB given, out = ∂f/∂τ at fixed `in`.
// B, dB, Cs, in, out are LDS arrays (in and out distinct); lane 0 writes the border rows, chain
// lanes their chains.  PREC: every lane loads its chain's inputs and coefficients before the border
// block, so their latency overlaps it (a lane reads SBZ_MCHAIN members from its chain's base,
// clamped to the state; the arithmetic is the same).
template <bool PREC = false>
SBZ_DEV void sbz_wave_apply(const WaveCtx& c, const double* B, const double* dB, const double* Cs, double k,
                            const double* in, double* out) {
    if constexpr (PREC) {
        double cv[SBZ_MCHAIN];
#pragma unroll
        for (int e = 0; e < SBZ_MCHAIN; ++e) cv[e] = in[(c.cbase + e < SBZ_N) ? c.cbase + e : SBZ_N - 1];
        const double att = in[c.cattc];
        const SbzChainRow r = dB ? sbz_chain_drow(c, B, dB, Cs, k) : sbz_chain_row(c, B, Cs, k);
        if (dB) sbz_wb_dfdt(c.lane, B, dB, Cs, k, in, out);
        else sbz_wb_f(c.lane, B, Cs, k, in, out);
        if (c.isC) {
#define SBZ_CG_ROWS_R(lo, hi, len) \
            if (SBZ_IN_GROUP(lo, hi)) sbz_chain_rows_r<len>(r, att, cv, out + c.cbase);
            SBZ_CHAIN_GROUPS(SBZ_CG_ROWS_R)
#undef SBZ_CG_ROWS_R
        }
        return;
    }
    if (dB) sbz_wb_dfdt(c.lane, B, dB, Cs, k, in, out);
    else sbz_wb_f(c.lane, B, Cs, k, in, out);
    if (c.isC) {
        const SbzChainRow r = dB ? sbz_chain_drow(c, B, dB, Cs, k) : sbz_chain_row(c, B, Cs, k);
        const double att = in[c.cattc];
#define SBZ_CG_ROWS(lo, hi, len) \
        if (SBZ_IN_GROUP(lo, hi)) sbz_chain_rows<len>(r, att, in + c.cbase, out + c.cbase);
        SBZ_CHAIN_GROUPS(SBZ_CG_ROWS)
#undef SBZ_CG_ROWS
    }
}


// Rodas5P tableau as compile-time constants, so the stage combinations below fold them into
// immediates (sbz_rodas5p.h).
struct SbzR5 {
    static constexpr double A[9][8] = {
        {0}, {0}, {SBZ_R5_A21}, {SBZ_R5_A31, SBZ_R5_A32}, {SBZ_R5_A41, SBZ_R5_A42, SBZ_R5_A43},
        {SBZ_R5_A51, SBZ_R5_A52, SBZ_R5_A53, SBZ_R5_A54},
        {SBZ_R5_A61, SBZ_R5_A62, SBZ_R5_A63, SBZ_R5_A64, SBZ_R5_A65},
        {SBZ_R5_A61, SBZ_R5_A62, SBZ_R5_A63, SBZ_R5_A64, SBZ_R5_A65, 1.0},
        {SBZ_R5_A61, SBZ_R5_A62, SBZ_R5_A63, SBZ_R5_A64, SBZ_R5_A65, 1.0, 1.0}};
    static constexpr double C[9][8] = {
        {0}, {0}, {SBZ_R5_C21}, {SBZ_R5_C31, SBZ_R5_C32}, {SBZ_R5_C41, SBZ_R5_C42, SBZ_R5_C43},
        {SBZ_R5_C51, SBZ_R5_C52, SBZ_R5_C53, SBZ_R5_C54},
        {SBZ_R5_C61, SBZ_R5_C62, SBZ_R5_C63, SBZ_R5_C64, SBZ_R5_C65},
        {SBZ_R5_C71, SBZ_R5_C72, SBZ_R5_C73, SBZ_R5_C74, SBZ_R5_C75, SBZ_R5_C76},
        {SBZ_R5_C81, SBZ_R5_C82, SBZ_R5_C83, SBZ_R5_C84, SBZ_R5_C85, SBZ_R5_C86, SBZ_R5_C87}};
    static constexpr double d[9] = {0, SBZ_R5_D1, SBZ_R5_D2, SBZ_R5_D3, SBZ_R5_D4, SBZ_R5_D5, 0, 0, 0};
};

// Stage combinations over the components a lane owns: i = lane + LN r for a mode spread over LN
// lanes (64: one mode per wavefront; 16: four modes per wavefront, lane = the lane within its row).
// stage S input  US = u_n + Σ_{j<S} a_Sj K_j  for this lane's components
template <int S, int LN>
SBZ_DEV void sbz_stage_in(int lane, const double (*v)[SBZ_N], const double* Un, double* US) {
#pragma unroll
    for (int r = 0; r < (SBZ_N + LN - 1) / LN; ++r) {
        const int i = lane + LN * r;
        if (i < SBZ_N) {
            double acc = Un[i];
#pragma unroll
            for (int j = 1; j < S; ++j) acc += SbzR5::A[S][j - 1] * v[5 + j][i];
            US[i] = acc;
        }
    }
}
// stage S right-hand side  K_S = f + h d_S ∂f/∂τ + Σ_{j<S} (c_Sj / h) K_j  (in place on K_S)
template <int S, int LN>
SBZ_DEV void sbz_stage_rhs(int lane, const double (*v)[SBZ_N], double dt, double idt, const double* DT, double* Ks) {
#pragma unroll
    for (int r = 0; r < (SBZ_N + LN - 1) / LN; ++r) {
        const int i = lane + LN * r;
        if (i < SBZ_N) {
            double acc = Ks[i] + (dt * SbzR5::d[S]) * DT[i];
#pragma unroll
            for (int j = 1; j < S; ++j) acc += (SbzR5::C[S][j - 1] * idt) * v[5 + j][i];
            Ks[i] = acc;
        }
    }
}
// The same two combinations with one read of each stage vector: the input pass also forms this
// lane's  ck = Σ_{j<S} (c_Sj / h) K_j  and keeps it in registers across the RHS application (for
// the one-wavefront-per-SIMD kernels; at two per SIMD the extra live registers cost more spills)
// With SBZ_STAGE_PRELOAD every load of a pass is unconditional (a lane past the state reads
// component N - 1 and discards it), so all of them issue together ahead of the arithmetic and only
// the stores are masked; without it each component's loads sit behind its bounds check, holding
// fewer registers (the mixed-model kernel with the several-modes bodies, short of them, takes that
// form).  The arithmetic is the same either way.
#ifndef SBZ_STAGE_PRELOAD
#define SBZ_STAGE_PRELOAD 1
#endif
template <int S, int LN>
SBZ_DEV void sbz_stage_in_f(int lane, const double (*v)[SBZ_N], const double* Un, double* US, double idt, double* ck) {
    if constexpr (!SBZ_STAGE_PRELOAD) {
#pragma unroll
        for (int r = 0; r < (SBZ_N + LN - 1) / LN; ++r) {
            const int i = lane + LN * r;
            if (i < SBZ_N) {
                double acc = Un[i], cc = 0.0;
#pragma unroll
                for (int j = 1; j < S; ++j) {
                    const double kj = v[5 + j][i];
                    acc += SbzR5::A[S][j - 1] * kj;
                    cc += (SbzR5::C[S][j - 1] * idt) * kj;
                }
                US[i] = acc;
                ck[r] = cc;
            }
        }
        return;
    }
    constexpr int OWNL = (SBZ_N + LN - 1) / LN;
    double un[OWNL], kv[OWNL][S > 1 ? S - 1 : 1];
#pragma unroll
    for (int r = 0; r < OWNL; ++r) {
        const int i = (lane + LN * r < SBZ_N) ? lane + LN * r : SBZ_N - 1;
        un[r] = Un[i];
#pragma unroll
        for (int j = 1; j < S; ++j) kv[r][j - 1] = v[5 + j][i];
    }
#pragma unroll
    for (int r = 0; r < OWNL; ++r) {
        const int i = lane + LN * r;
        double acc = un[r], cc = 0.0;
#pragma unroll
        for (int j = 1; j < S; ++j) {
            acc += SbzR5::A[S][j - 1] * kv[r][j - 1];
            cc += (SbzR5::C[S][j - 1] * idt) * kv[r][j - 1];
        }
        if (i < SBZ_N) US[i] = acc;
        ck[r] = cc;
    }
}
template <int S, int LN>
SBZ_DEV void sbz_stage_rhs_f(int lane, double dt, const double* DT, double* Ks, const double* ck) {
    if constexpr (!SBZ_STAGE_PRELOAD) {
#pragma unroll
        for (int r = 0; r < (SBZ_N + LN - 1) / LN; ++r) {
            const int i = lane + LN * r;
            if (i < SBZ_N) Ks[i] = (Ks[i] + (dt * SbzR5::d[S]) * DT[i]) + ck[r];
        }
        return;
    }
    constexpr int OWNL = (SBZ_N + LN - 1) / LN;
    double kk[OWNL], dd[OWNL];
#pragma unroll
    for (int r = 0; r < OWNL; ++r) {
        const int i = (lane + LN * r < SBZ_N) ? lane + LN * r : SBZ_N - 1;
        kk[r] = Ks[i];
        dd[r] = DT[i];
    }
#pragma unroll
    for (int r = 0; r < OWNL; ++r) {
        const int i = lane + LN * r;
        if (i < SBZ_N) Ks[i] = (kk[r] + (dt * SbzR5::d[S]) * dd[r]) + ck[r];
    }
}
#define SBZ_STAGE_SWITCH_LN(F, LN, ...)                                                           \
    switch (s) {                                                                                  \
    case 2: F<2, LN>(__VA_ARGS__); break;                                                         \
    case 3: F<3, LN>(__VA_ARGS__); break;                                                         \
    case 4: F<4, LN>(__VA_ARGS__); break;                                                         \
    case 5: F<5, LN>(__VA_ARGS__); break;                                                         \
    case 6: F<6, LN>(__VA_ARGS__); break;                                                         \
    case 7: F<7, LN>(__VA_ARGS__); break;                                                         \
    default: F<8, LN>(__VA_ARGS__); break;                                                        \
    }
#define SBZ_STAGE_SWITCH(F, ...) SBZ_STAGE_SWITCH_LN(F, 64, __VA_ARGS__)

// Wave-wide sum, identical in every lane (called with all 64 lanes active): butterflies within
// each row of 16 lanes (quad_perm xor 1, xor 2, row_half_mirror, row_mirror), then the four row
// sums combined through readlane — every control decision taken on it must be uniform.
SBZ_DEV double sbz_wave_sum(double v) {
    v += sbz_dpp<0xB1>(v);    // quad_perm [1,0,3,2]
    v += sbz_dpp<0x4E>(v);    // quad_perm [2,3,0,1]
    v += sbz_dpp<0x141>(v);   // row_half_mirror
    v += sbz_dpp<0x140>(v);   // row_mirror
    return (sbz_readlane(v, 0) + sbz_readlane(v, 16)) + (sbz_readlane(v, 32) + sbz_readlane(v, 48));
}

// The lane roles of a 64-lane wavefront that integrates one mode.
SBZ_DEV WaveCtx sbz_wave_ctx(int lane) {
    WaveCtx c;
    c.lane = lane;
    c.isB = c.lane < SBZ_NBORD;
    c.bcomp = c.isB ? sbz_bord_comp[c.lane] : 0;
    c.bchain = c.isB ? sbz_bord_chain[c.lane] : -1;
    c.isC = c.lane < SBZ_NCHAIN;
    const int ci = c.isC ? c.lane : 0;
    c.ckind = sbz_chain_kind[ci];
    c.cL = sbz_chain_L[ci];
    c.cbin = sbz_chain_bin[ci];
    c.catt = sbz_chain_att[ci];
    c.cattc = sbz_chain_attc[ci];
    c.cbase = sbz_chain_base[ci];
    return c;
}

// Integrate instance `inst` from its start to τ0 (saves and sources on the way) and write its
// final state, return code and statistics; every lane of the wavefront calls it.
template <bool FUSED, bool SPLIT = SBZ_LU_SPLIT>
SBZ_DEV void sbz_wave_mode(const SbzSolveArgs& a, WaveLDSt<SPLIT>& L, const WaveCtx& c, int inst) {
    // the per-attempt passes over the components issue all their loads first (the fused one-wavefront-
    // per-SIMD kernels, with registers to spare)
    constexpr bool LOADS_FIRST = FUSED && SBZ_STAGE_PRELOAD;
    double* const DT = L.v[4];
    double* const US = L.v[5];
    double* const TMP = L.v[8];   // source scratch: K3 is dead between step attempts (rewritten at stage 3)
    {
        const int set = a.inst_set[inst];
        const double* saveat = a.saveat ? a.saveat + (size_t)set * a.saveat_stride : nullptr;
        const double* P = a.params + (size_t)set * SBZ_NPAR;
        const double* kt = a.knots + a.knot_off[set];
        const int nk = a.knot_n[set];
        const double k = a.inst_k[inst];
        const double t1 = kt[nk - 1];
        double t = a.inst_t0[inst];
        int kidx = sbz_find_knot(kt, nk, t);
        int ub = 0, fb = 2;  // ping-pong slots of u_n and f_n
        long long st_acc = 0, st_rej = 0, st_f = 0, st_j = 0;
        int ret = SBZ_RET_SUCCESS;
        int isave = 0;
        double* osave = a.out_save ? a.out_save + (size_t)inst * a.n_save * SBZ_N : nullptr;
        double* osrc = a.out_src ? a.out_src + (size_t)inst * a.n_save * 3 : nullptr;
        if (c.lane == 0) {
            double Cs[SBZ_NC], Bx[SBZ_NB];
            sbz_consts(P, Cs);
            sbz_consts_k(Cs, k);
#pragma unroll
            for (int q = 0; q < SBZ_NC; ++q) L.Cs[q] = Cs[q];
            int kk = kidx;
            sbz_bg_plain(P, kt, nk, kk, t, Bx);
#pragma unroll
            for (int q = 0; q < SBZ_NB; ++q) L.B[6][q] = Bx[q];
            sbz_ic(L.B[6], Cs, k, L.v[0]);
        }
        __syncthreads();
        {
            bool fin = (k > 0.0) && (k == k);
            for (int i = c.lane; i < SBZ_N; i += 64) {
                const double v = L.v[0][i];
                if (!(v == v) || fabs(v) > 1e300) fin = false;
            }
            if (!__all(fin)) ret = SBZ_RET_INITIALFAILURE;
        }
        for (int i = c.lane; i < SBZ_N; i += 64)
#pragma unroll
            for (int j = 6; j <= 13; ++j) L.v[j][i] = 0.0;   // stage vectors start finite (LDS is not initialised)
        double dt = 0.0, qold = SBZ_CTL_QOLDINIT;
        if (ret == SBZ_RET_SUCCESS && t1 > t) {
            // Hairer–Wanner initial step: f(t, u0) and f(t + dt0, u0 + dt0 f0)
            double d0 = 0.0, d1 = 0.0, dt0 = 0.0;
            for (int pass = 0; pass < 2; ++pass) {
                if (pass == 1 && c.lane == 0) {
                    double Bx[SBZ_NB];
                    int kk = kidx;
                    sbz_bg_plain(P, kt, nk, kk, t + dt0, Bx);
#pragma unroll
                    for (int q = 0; q < SBZ_NB; ++q) L.B[6][q] = Bx[q];
                }
                __syncthreads();
                sbz_wave_apply(c, L.B[6], nullptr, L.Cs, k, L.v[pass], L.v[2 + pass]);
                __syncthreads();
                if (pass == 0) {
                    for (int i = c.lane; i < SBZ_N; i += 64) {
                        const double sk = a.atol + fabs(L.v[0][i]) * a.rtol;
                        d0 += (L.v[0][i] / sk) * (L.v[0][i] / sk);
                        d1 += (L.v[2][i] / sk) * (L.v[2][i] / sk);
                    }
                    d0 = sqrt(sbz_wave_sum(d0) / SBZ_N);
                    d1 = sqrt(sbz_wave_sum(d1) / SBZ_N);
                    dt0 = (d0 < 1e-5 || d1 < 1e-5) ? 1e-6 : 0.01 * d0 / d1;
                    dt0 = fmin(dt0, t1 - t);
                    for (int i = c.lane; i < SBZ_N; i += 64) L.v[1][i] = L.v[0][i] + dt0 * L.v[2][i];
                }
            }
            double d2 = 0.0;
            for (int i = c.lane; i < SBZ_N; i += 64) {
                const double sk = a.atol + fabs(L.v[0][i]) * a.rtol;
                const double r = (L.v[3][i] - L.v[2][i]) / sk;
                d2 += r * r;
            }
            d2 = sqrt(sbz_wave_sum(d2) / SBZ_N) / dt0;
            const double mx = fmax(d1, d2);
            const double dt1 = (mx <= 1e-15) ? fmax(1e-6, dt0 * 1e-3) : pow(0.01 / mx, 1.0 / 6.0);
            dt = fmin(fmin(100.0 * dt0, dt1), t1 - t);
            st_f += 2;
        }
        double row[SBZ_NBORD];
        double cpl = 0.0;
        SbzWb wb;                  // low-rank (Woodbury) data of the border solve
        double cinv[SBZ_MCHAIN];   // reciprocal modified pivots of this lane's chain
        double csk = 0.0;          // s k of this lane's chain at t_n
        int bidx = kidx;   // knot cursor of this lane's stage-time background
        long long iters = 0;
        double tprev = t, hprev = 0.0;  // the last accepted step, for dense output
        bool have_step = false;
        for (;;) {
            // ---- saves in (tprev, t] (or at/before the start): Hermite dense output, sources
            while ((osave || osrc) && isave < a.n_save && saveat[isave] <= t) {
                const int un = ub, u0 = ub ^ 1, fn = fb, f0 = fb ^ 1;  // current (n) and previous slots
                const double th = have_step ? (saveat[isave] - tprev) / hprev : 1.0;
                const double h00 = (1 + 2 * th) * (1 - th) * (1 - th), h10 = th * (1 - th) * (1 - th);
                const double h01 = th * th * (3 - 2 * th), h11 = th * th * (th - 1);
                for (int i = c.lane; i < SBZ_N; i += 64) {
                    const double y = have_step ? h00 * L.v[u0][i] + h10 * hprev * L.v[f0][i] + h01 * L.v[un][i] +
                                                     h11 * hprev * L.v[fn][i]
                                               : L.v[un][i];
                    US[i] = y;
                    if (osave) osave[(size_t)isave * SBZ_N + i] = y;
                }
                if (osrc && ret == SBZ_RET_SUCCESS) {
                    if (a.srcbg) {   // the set's table row (sbz_srcbg_kernel), spread over the lanes
                        const double* sb = a.srcbg + ((size_t)set * a.n_save + isave) * SBZ_SRCBG_STRIDE;
                        for (int q = c.lane; q <= 2 * SBZ_NB; q += 64) {
                            const double v = sb[q];
                            if (q < SBZ_NB) L.B[6][q] = v;
                            else if (q < 2 * SBZ_NB) L.dB[1][q - SBZ_NB] = v;
                            else L.ddkd = v;
                        }
                    } else if (c.lane == 0) {
                        double Bt[SBZ_NB], dBt[SBZ_NB];
                        int kk = kidx;
                        L.ddkd = sbz_bg_tay2(P, kt, nk, kk, saveat[isave], Bt, dBt);
#pragma unroll
                        for (int q = 0; q < SBZ_NB; ++q) { L.B[6][q] = Bt[q]; L.dB[1][q] = dBt[q]; }
                    }
                    __syncthreads();
                    // u' = f(u), J u' = f(u'), ∂f/∂τ (u), u'' = ∂f/∂τ + J u'
                    for (int pass = 0; pass < 3; ++pass) {
                        const double* in = (pass == 1) ? L.v[6] : US;
                        double* out = (pass == 0) ? L.v[6] : (pass == 1) ? L.v[7] : TMP;
                        sbz_wave_apply(c, L.B[6], pass == 2 ? L.dB[1] : nullptr, L.Cs, k, in, out);
                        __syncthreads();
                    }
                    for (int i = c.lane; i < SBZ_N; i += 64) L.v[7][i] += TMP[i];
                    __syncthreads();
                    double S3[3];
                    sbz_src(L.B[6], L.dB[1], L.ddkd, L.Cs, k, P[SBZ_P_TAU0], P[SBZ_P_TAUREC], US, L.v[6], L.v[7], S3);
                    if (c.lane == 0) {
                        osrc[(size_t)isave * 3 + 0] = S3[0];
                        osrc[(size_t)isave * 3 + 1] = S3[1];
                        osrc[(size_t)isave * 3 + 2] = S3[2];
                    }
                }
                __syncthreads();
                ++isave;
            }
            if (!(ret == SBZ_RET_SUCCESS && t < t1)) break;
            // ---- one Rodas5P step attempt from (t, u_n)
            double* const Un = L.v[ub];
            double* const U1 = L.v[ub ^ 1];
            double* const F0 = L.v[fb];
            double* const F1 = L.v[fb ^ 1];
            if (++iters > a.maxiters) { ret = SBZ_RET_MAXITERS; break; }
            if (t + dt >= t1 - 1e-12 * fabs(t1)) dt = t1 - t;
            else if (dt < 1e-14 * fabs(t)) { ret = SBZ_RET_DTLESSTHANMIN; break; }
            const double ig = 1.0 / (SBZ_R5_GAMMA * dt);
            const double idt = 1.0 / dt;
            // background at the six distinct stage times, one lane each (dual: d/dτ from lane 0)
            if (c.lane < 6) {
                const double cl = (c.lane == 0) ? 0.0 : (c.lane == 1) ? SBZ_R5_C2 : (c.lane == 2) ? SBZ_R5_C3
                                : (c.lane == 3) ? SBZ_R5_C4 : (c.lane == 4) ? SBZ_R5_C5 : 1.0;
                double Bq[SBZ_NB], dBq[SBZ_NB];
                sbz_bg_dual(P, kt, nk, bidx, t + cl * dt, Bq, dBq);
#pragma unroll
                for (int q = 0; q < SBZ_NB; ++q) L.B[c.lane][q] = Bq[q];
                if (c.lane == 0) {
#pragma unroll
                    for (int q = 0; q < SBZ_NB; ++q) L.dB[0][q] = dBq[q];
                }
            }
            __syncthreads();
            const double* B0 = L.B[0];
            // chain factorization (top-down elimination from l = L) and its Schur term
            if (c.isC) {
                const SbzChainRow r = sbz_chain_row(c, B0, L.Cs, k);
                csk = r.sk;
                double xcs = 0.0;
#define SBZ_CG_FACTOR(lo, hi, len) \
                if (SBZ_IN_GROUP(lo, hi)) xcs = sbz_chain_factor<len>(r, ig, cinv);
                SBZ_CHAIN_GROUPS(SBZ_CG_FACTOR)
#undef SBZ_CG_FACTOR
                L.C[c.lane] = xcs;
            }
            __syncthreads();
            const double xc = (c.isB && c.bchain >= 0) ? L.C[c.bchain] : 0.0;
            double al, be;
            sbz_wb_w(c.lane, B0, L.Cs, k, ig, xc, sbz_wave_m0(L), row, cpl, al, be, wb);
            double rpo;   // reciprocal pivot of this lane's row
            bool ok = sbz_wb_lu<SbzBcWave>(c.lane, row, rpo);
            sbz_wb_store<SPLIT>(c.lane, row, rpo, L.LU);
            // Woodbury data: z = A^-1 [α β] and C^-1 with C = I - [p q]^T z (2 x 2)
            sbz_wb_tri2<SPLIT, SbzBcWave>(c.lane, L.LU, al, be);
            wb.z1 = c.isB ? al : 0.0;
            wb.z2 = c.isB ? be : 0.0;
            {
                const double c11 = 1.0 - sbz_wave_sum(wb.p * wb.z1), c12 = -sbz_wave_sum(wb.p * wb.z2);
                const double c21 = -sbz_wave_sum(wb.q * wb.z1), c22 = 1.0 - sbz_wave_sum(wb.q * wb.z2);
                const double det = c11 * c22 - c12 * c21;
                ok = ok && (det != 0.0) && (det == det) && (fabs(det) < 1e300);
                const double idet = 1.0 / det;
                wb.c11 = c22 * idet;
                wb.c12 = -c12 * idet;
                wb.c21 = -c21 * idet;
                wb.c22 = c11 * idet;
            }
            ok = __all(ok || !c.isB);
            sbz_wave_apply(c, B0, L.dB[0], L.Cs, k, Un, DT);  // ∂f/∂τ at (t, u_n)
            st_j++;
            __syncthreads();
            double EEst = __builtin_inf();
            if (ok) {
                // stages 1..8, then pass 9 evaluates f(t + h, u_{n+1}) for the next step / dense output
                for (int s = 1; s <= 9; ++s) {
                    double* Ks = (s <= 8) ? L.v[5 + s] : F1;
                    if (s == 1) {
                        if constexpr (LOADS_FIRST) {   // every load first, masked stores
                            double f0[SBZ_OWN], d0[SBZ_OWN];
#pragma unroll
                            for (int r = 0; r < SBZ_OWN; ++r) {
                                const int i = (c.lane + 64 * r < SBZ_N) ? c.lane + 64 * r : SBZ_N - 1;
                                f0[r] = F0[i];
                                d0[r] = DT[i];
                            }
#pragma unroll
                            for (int r = 0; r < SBZ_OWN; ++r)
                                if (c.lane + 64 * r < SBZ_N) Ks[c.lane + 64 * r] = f0[r] + dt * SBZ_R5_D1 * d0[r];
                        } else {
                            for (int i = c.lane; i < SBZ_N; i += 64) Ks[i] = F0[i] + dt * SBZ_R5_D1 * DT[i];
                        }
                    } else {
                        double* in = (s <= 8) ? US : U1;
                        double ck[SBZ_OWN];   // fused: Σ_{j<s} (c_sj / h) K_j of this lane's components
                        if (s <= 8) {
                            if constexpr (FUSED) SBZ_STAGE_SWITCH(sbz_stage_in_f, c.lane, L.v, Un, US, idt, ck)
                            else SBZ_STAGE_SWITCH(sbz_stage_in, c.lane, L.v, Un, US)
                        }
                        __syncthreads();
                        sbz_wave_apply(c, L.B[sbz_stage_bg[s]], nullptr, L.Cs, k, in, Ks);
                        st_f++;
                        __syncthreads();
                        if (s == 9) break;
                        if constexpr (FUSED) SBZ_STAGE_SWITCH(sbz_stage_rhs_f, c.lane, dt, DT, Ks, ck)
                        else SBZ_STAGE_SWITCH(sbz_stage_rhs, c.lane, L.v, dt, idt, DT, Ks)
                    }
                    __syncthreads();
                    // solve W x = Ks: chain forward sweeps, border solve, chain back-substitution
                    // chain forward sweeps (in place); the attachment rows take their Schur terms
                    if (c.isC) {
                        double y = 0.0;
#define SBZ_CG_FWD(lo, hi, len) \
                        if (SBZ_IN_GROUP(lo, hi)) y = sbz_chain_fwd<len>(csk, cinv, Ks + c.cbase);
                        SBZ_CHAIN_GROUPS(SBZ_CG_FWD)
#undef SBZ_CG_FWD
                        L.C[c.lane] = y;
                    }
                    __syncthreads();
                    {
                        double b = c.isB ? Ks[c.bcomp] : 0.0;
                        if (c.isB && c.bchain >= 0) b -= cpl * L.C[c.bchain];
                        sbz_wb_tri<SPLIT, SbzBcWave>(c.lane, L.LU, b);   // y = A^-1 b
                        const double s1 = sbz_wave_sum(wb.p * b), s2 = sbz_wave_sum(wb.q * b);
                        const double x = b + wb.z1 * (wb.c11 * s1 + wb.c12 * s2) + wb.z2 * (wb.c21 * s1 + wb.c22 * s2);
                        if (c.isB) {
                            Ks[c.bcomp] = x;
                            L.X[c.lane] = x;
                        }
                    }
                    __syncthreads();
                    // chain back substitutions from the solved attachments
                    if (c.isC) {
                        const double xatt = L.X[c.catt];
#define SBZ_CG_BWD(lo, hi, len) \
                        if (SBZ_IN_GROUP(lo, hi)) sbz_chain_bwd<len>(csk, cinv, xatt, Ks + c.cbase);
                        SBZ_CHAIN_GROUPS(SBZ_CG_BWD)
#undef SBZ_CG_BWD
                    }
                    __syncthreads();
                    if (s == 8) {
                        // u_{n+1} = stage-8 input + k8; error estimate k8
                        double ss = 0.0;
                        bool fin = true;
                        if constexpr (LOADS_FIRST) {   // every load first, as for stage 1
                            double us[SBZ_OWN], kv[SBZ_OWN], uo[SBZ_OWN];
#pragma unroll
                            for (int r = 0; r < SBZ_OWN; ++r) {
                                const int i = (c.lane + 64 * r < SBZ_N) ? c.lane + 64 * r : SBZ_N - 1;
                                us[r] = US[i];
                                kv[r] = Ks[i];
                                uo[r] = Un[i];
                            }
#pragma unroll
                            for (int r = 0; r < SBZ_OWN; ++r) {
                                const int i = c.lane + 64 * r;
                                if (i < SBZ_N) {
                                    const double un1 = us[r] + kv[r];
                                    U1[i] = un1;
                                    const double sc = a.atol + fmax(fabs(uo[r]), fabs(un1)) * a.rtol;
                                    const double q = kv[r] / sc;
                                    ss += q * q;
                                    if (!(un1 == un1) || fabs(un1) > 1e300) fin = false;
                                }
                            }
                        } else {
                            for (int i = c.lane; i < SBZ_N; i += 64) {
                                const double un1 = US[i] + Ks[i];
                                U1[i] = un1;
                                const double sc = a.atol + fmax(fabs(Un[i]), fabs(un1)) * a.rtol;
                                const double q = Ks[i] / sc;
                                ss += q * q;
                                if (!(un1 == un1) || fabs(un1) > 1e300) fin = false;
                            }
                        }
                        ss = sbz_wave_sum(ss);
                        EEst = __all(fin) ? sqrt(ss / SBZ_N) : __builtin_inf();
                        if (!(EEst <= 1.0)) break;   // rejected: no f(t + h) needed
                    }
                }
            }
            __syncthreads();
            // the controller's two powers EEst^β1 and qold^β2: in the one-wavefront kernel as one
            // evaluation on lanes 0 and 1; the two-per-SIMD build keeps two calls, which spill less there
            double q11, qob = 1.0;
            if constexpr (FUSED) {
                const double pw = pow((c.lane == 1) ? qold : ((EEst < __builtin_inf()) ? EEst : 1.0),
                                      (c.lane == 1) ? SBZ_CTL_BETA2 : SBZ_CTL_BETA1);
                q11 = sbz_readlane(pw, 0);
                qob = sbz_readlane(pw, 1);
            } else {
                q11 = (EEst < __builtin_inf()) ? pow(EEst, SBZ_CTL_BETA1) : 1.0;
            }
            if (!(EEst <= 1.0)) {
                const double q11r = (EEst < __builtin_inf()) ? q11 : 1.0 / SBZ_CTL_QMIN;
                dt /= fmin(1.0 / SBZ_CTL_QMIN, q11r / SBZ_CTL_GAMMA);
                st_rej++;
                // the stage combinations read every stage vector with zero weights past the
                // stage, so a non-finite attempt must not leave non-finite stages behind
                if (!(EEst < __builtin_inf())) {
                    for (int i = c.lane; i < SBZ_N; i += 64)
#pragma unroll
                        for (int j = 6; j <= 13; ++j) L.v[j][i] = 0.0;
                    __syncthreads();
                }
                continue;
            }
            const double q = (EEst == 0.0) ? 1.0 / SBZ_CTL_QMAX
                                           : fmax(1.0 / SBZ_CTL_QMAX,
                                                  fmin(1.0 / SBZ_CTL_QMIN,
                                                       q11 / (FUSED ? qob : pow(qold, SBZ_CTL_BETA2)) / SBZ_CTL_GAMMA));
            qold = fmax(EEst, SBZ_CTL_QOLDINIT);
            double tn = t + dt;
            if (tn >= t1 - 1e-14 * fabs(t1)) tn = t1;
            tprev = t;
            hprev = dt;
            have_step = true;
            t = tn;
            ub ^= 1;
            fb ^= 1;
            while (kidx < nk - 2 && t >= kt[kidx + 1]) ++kidx;
            st_acc++;
            dt = dt / q;
        }
        __syncthreads();
        {
            double* o = a.out_u + (size_t)inst * SBZ_N;
            for (int i = c.lane; i < SBZ_N; i += 64) o[i] = L.v[ub][i];
        }
        if (c.lane == 0) {
            a.retcode[inst] = ret;
            long long* st = a.stats + (size_t)inst * 6;
            st[0] = st_acc;
            st[1] = st_rej;
            st[2] = st_f;
            st[3] = st_j;
            st[4] = st_j;
            st[5] = 8 * st_j;
        }
        __syncthreads();
    }
}

template <bool FUSED, bool SPLIT = SBZ_LU_SPLIT>
SBZ_DEV void sbz_wave_body(const SbzSolveArgs& a) {
    __shared__ WaveLDSt<SPLIT> L;
    const WaveCtx c = sbz_wave_ctx(threadIdx.x);
    for (;;) {
        if (c.lane == 0) L.mode = sbz_pull(a, (int)(blockIdx.x % (unsigned)a.n_xcd));
        __syncthreads();
        const int qi = L.mode;
        __syncthreads();
        if (qi < 0) break;
        sbz_wave_mode<FUSED, SPLIT>(a, L, c, a.order ? a.order[qi] : qi);
    }
}

#if SBZ_WAVE4_OK || SBZ_WAVE2H_OK
// ============================================================================================
// Several modes per wavefront: the LN lanes of row g = lane / LN integrate one mode, for models
// whose border and chains fit LN lanes.  LN = 16 (SBZ_ALGO_WAVE4, four modes per wavefront) or
// LN = 32 (SBZ_ALGO_WAVE2H, two modes per wavefront).  Per mode the work is sbz_wave_body's, with
// the wavefront broadcasts and reductions in their row forms (16 lanes: DPP row_newbcast and a
// row butterfly; 32 lanes: readlane per half and the row butterfly plus a swap of the half's two
// 16-lane rows).  The rows share one uniform control flow: every iteration each row takes one
// step attempt; accept/reject, step size and time are per row (uniform within the row); a row
// whose mode ends writes it and pulls the next instance.  Rows without a mode (the queue is
// empty) step on harmless values and discard them.  States only: no saves, no sources.
// ============================================================================================
struct WaveRow {
    double v[14][SBZ_N];           // 0,1 U (ping-pong), 2,3 F (ping-pong), 4 DT, 5 US, 6..13 K1..K8
    double B[6][SBZ_NB];           // background at the six distinct stage times
    double dB[SBZ_NB];             // d/dτ of the background at t_n
    double Cs[SBZ_NC];             // per-mode constants
    double M[SBZ_NBORD * SBZ_WSTRIDE];     // border block staging (the factors stay in registers)
    double X[SBZ_NBORD];           // border solution broadcast
    double C[SBZ_NCHAIN];          // chain → border exchange
    int mode;
};

// Sum over the 16 lanes of each row, returned in every lane of the row (all 64 lanes active).
SBZ_DEV double sbz_row_sum(double v) {
    v += sbz_dpp<0xB1>(v);    // quad_perm [1,0,3,2]
    v += sbz_dpp<0x4E>(v);    // quad_perm [2,3,0,1]
    v += sbz_dpp<0x141>(v);   // row_half_mirror
    v += sbz_dpp<0x140>(v);   // row_mirror
    return v;
}
// The row operations of LN-lane rows (all 64 lanes active): the broadcast class of the border
// solves, the sum over the row (identical in every lane of the row) and whether p holds in every
// lane of row g.
template <int LN>
struct SbzRows;
template <>
struct SbzRows<16> {
    using BC = SbzBcRow16;
    static SBZ_DEV double sum(double v) { return sbz_row_sum(v); }
    static SBZ_DEV bool all(bool p, int g) {
        const unsigned long long m = __ballot(p);
        return ((m >> (16 * g)) & 0xFFFFull) == 0xFFFFull;
    }
};
template <>
struct SbzRows<32> {
    using BC = SbzBcHalf32;
    static SBZ_DEV double sum(double v) {
        v = sbz_row_sum(v);
        return v + __shfl_xor(v, 16, 32);   // the other 16-lane row of the half
    }
    static SBZ_DEV bool all(bool p, int g) {
        const unsigned long long m = __ballot(p);
        return ((m >> (32 * g)) & 0xFFFFFFFFull) == 0xFFFFFFFFull;
    }
};

// The rows of one wavefront integrate instances pulled from a's queue until it is empty; R4 holds
// the 64 / LN rows' LDS.
template <bool FUSED, int LN>
SBZ_DEV void sbz_waverow_run(const SbzSolveArgs& a, WaveRow* R4) {
    using RO = SbzRows<LN>;
    using BC = typename RO::BC;
    constexpr int OWN = (SBZ_N + LN - 1) / LN;
    // the per-attempt passes over a row's components issue all their loads first at 16-lane rows
    // (registers to spare); 32-lane rows and the mixed kernel keep the bounds-checked loops
    constexpr bool LOADS_FIRST = SBZ_STAGE_PRELOAD && LN == 16;
    const int g = threadIdx.x / LN;
    WaveRow& L = R4[g];
    WaveCtx c;
    c.lane = threadIdx.x & (LN - 1);
    c.isB = c.lane < SBZ_NBORD;
    c.bcomp = c.isB ? sbz_bord_comp[c.lane] : 0;
    c.bchain = c.isB ? sbz_bord_chain[c.lane] : -1;
    c.isC = c.lane < SBZ_NCHAIN;
    const int ci = c.isC ? c.lane : 0;
    c.ckind = sbz_chain_kind[ci];
    c.cL = sbz_chain_L[ci];
    c.cbin = sbz_chain_bin[ci];
    c.catt = sbz_chain_att[ci];
    c.cattc = sbz_chain_attc[ci];
    c.cbase = sbz_chain_base[ci];
    double* const DT = L.v[4];
    double* const US = L.v[5];
    // per-row state; a row without a mode keeps the tables of a valid set so its steps stay finite
    bool have = false, exhausted = false;
    int inst = 0;
    int set = a.inst_set[a.order ? a.order[0] : 0];
    const double* P = a.params + (size_t)set * SBZ_NPAR;
    const double* kt = a.knots + a.knot_off[set];
    int nk = a.knot_n[set];
    double k = 1.0, t = kt[0], t1 = kt[nk - 1], dt = 1e-3 * (kt[nk - 1] - kt[0]);
    int kidx = 0, bidx = 0, ub = 0, fb = 2;
    long long st_acc = 0, st_rej = 0, st_f = 0, st_j = 0, iters = 0;
    int ret = SBZ_RET_SUCCESS;
    double qold = SBZ_CTL_QOLDINIT;
    double row[SBZ_NBORD];
    double cpl = 0.0;
    SbzWb wb;
    double cinv[SBZ_MCHAIN];
    double csk = 0.0;
    // the backward border solves scale by the reciprocal pivots after their broadcast, from
    // registers (16-lane rows), or before it, in the pivot's own lane (32-lane rows, which are
    // short of registers)
    constexpr bool OWNSCALE = (LN == 32);
    double lrow[SBZ_NBORD], urow[SBZ_NBORD];   // this lane's factored border row
    double rpb[OWNSCALE ? 1 : SBZ_NBORD];      // the broadcast reciprocal pivots
    double rpo = 0.0;                           // its reciprocal pivot
    for (int i = c.lane; i < SBZ_N; i += LN)
#pragma unroll
        for (int j = 0; j < 14; ++j) L.v[j][i] = 0.0;
    if (c.lane == 0) {
        double Cs[SBZ_NC];
        sbz_consts(P, Cs);
        sbz_consts_k(Cs, k);
#pragma unroll
        for (int q = 0; q < SBZ_NC; ++q) L.Cs[q] = Cs[q];
    }
    __syncthreads();
    // the mode of this row ends: write its state, return code and statistics
    auto finish = [&]() {
        double* o = a.out_u + (size_t)inst * SBZ_N;
        for (int i = c.lane; i < SBZ_N; i += LN) o[i] = L.v[ub][i];
        if (c.lane == 0) {
            a.retcode[inst] = ret;
            long long* st = a.stats + (size_t)inst * 6;
            st[0] = st_acc;
            st[1] = st_rej;
            st[2] = st_f;
            st[3] = st_j;
            st[4] = st_j;
            st[5] = 8 * st_j;
        }
        have = false;
    };
    for (;;) {
        // ---- rows without a mode pull the next instance and start it
        const bool need = !have && !exhausted;
        if (__any(need)) {
            if (need && c.lane == 0) L.mode = sbz_pull(a, (int)(blockIdx.x % (unsigned)a.n_xcd));
            __syncthreads();
            const int qi = L.mode;
            __syncthreads();
            const bool fresh = need && qi >= 0;
            if (need && !fresh) exhausted = true;
            if (__any(fresh)) {
                if (fresh) {
                    inst = a.order ? a.order[qi] : qi;
                    set = a.inst_set[inst];
                    P = a.params + (size_t)set * SBZ_NPAR;
                    kt = a.knots + a.knot_off[set];
                    nk = a.knot_n[set];
                    k = a.inst_k[inst];
                    t1 = kt[nk - 1];
                    t = a.inst_t0[inst];
                    kidx = sbz_find_knot(kt, nk, t);
                    bidx = kidx;
                    ub = 0;
                    fb = 2;
                    st_acc = st_rej = st_f = st_j = iters = 0;
                    ret = SBZ_RET_SUCCESS;
                    qold = SBZ_CTL_QOLDINIT;
                    if (c.lane == 0) {
                        double Cs[SBZ_NC], Bx[SBZ_NB];
                        sbz_consts(P, Cs);
                        sbz_consts_k(Cs, k);
#pragma unroll
                        for (int q = 0; q < SBZ_NC; ++q) L.Cs[q] = Cs[q];
                        int kk = kidx;
                        sbz_bg_plain(P, kt, nk, kk, t, Bx);
#pragma unroll
                        for (int q = 0; q < SBZ_NB; ++q) L.B[0][q] = Bx[q];
                        sbz_ic(L.B[0], Cs, k, L.v[0]);
                    }
                }
                __syncthreads();
                {
                    bool fin = (k > 0.0) && (k == k);
                    for (int i = c.lane; i < SBZ_N; i += LN) {
                        const double v = L.v[0][i];
                        if (!(v == v) || fabs(v) > 1e300) fin = false;
                    }
                    const bool fin_row = RO::all(fin, g);
                    if (fresh && !fin_row) ret = SBZ_RET_INITIALFAILURE;
                }
                if (fresh)
                    for (int i = c.lane; i < SBZ_N; i += LN)
#pragma unroll
                        for (int j = 6; j <= 13; ++j) L.v[j][i] = 0.0;   // stage vectors start finite
                // Hairer–Wanner initial step of the starting rows; the other rows evaluate on their
                // (dead) stage vectors and discard the result
                const bool init = fresh && ret == SBZ_RET_SUCCESS && t1 > t;
                double d0 = 0.0, d1 = 0.0;
                sbz_wave_apply<LOADS_FIRST>(c, L.B[0], nullptr, L.Cs, k, init ? L.v[0] : L.v[6], init ? L.v[2] : L.v[7]);
                __syncthreads();
                if (init)
                    for (int i = c.lane; i < SBZ_N; i += LN) {
                        const double sk = a.atol + fabs(L.v[0][i]) * a.rtol;
                        d0 += (L.v[0][i] / sk) * (L.v[0][i] / sk);
                        d1 += (L.v[2][i] / sk) * (L.v[2][i] / sk);
                    }
                d0 = sqrt(RO::sum(d0) / SBZ_N);
                d1 = sqrt(RO::sum(d1) / SBZ_N);
                double dt0 = (d0 < 1e-5 || d1 < 1e-5) ? 1e-6 : 0.01 * d0 / d1;
                dt0 = fmin(dt0, t1 - t);
                if (init) {
                    for (int i = c.lane; i < SBZ_N; i += LN) L.v[1][i] = L.v[0][i] + dt0 * L.v[2][i];
                    if (c.lane == 0) {
                        double Bx[SBZ_NB];
                        int kk = kidx;
                        sbz_bg_plain(P, kt, nk, kk, t + dt0, Bx);
#pragma unroll
                        for (int q = 0; q < SBZ_NB; ++q) L.B[1][q] = Bx[q];
                    }
                }
                __syncthreads();
                sbz_wave_apply<LOADS_FIRST>(c, L.B[1], nullptr, L.Cs, k, init ? L.v[1] : L.v[6], init ? L.v[3] : L.v[7]);
                __syncthreads();
                double d2 = 0.0;
                if (init)
                    for (int i = c.lane; i < SBZ_N; i += LN) {
                        const double sk = a.atol + fabs(L.v[0][i]) * a.rtol;
                        const double r = (L.v[3][i] - L.v[2][i]) / sk;
                        d2 += r * r;
                    }
                d2 = sqrt(RO::sum(d2) / SBZ_N) / dt0;
                if (init) {
                    const double mx = fmax(d1, d2);
                    const double dt1 = (mx <= 1e-15) ? fmax(1e-6, dt0 * 1e-3) : pow(0.01 / mx, 1.0 / 6.0);
                    dt = fmin(fmin(100.0 * dt0, dt1), t1 - t);
                    st_f += 2;
                }
                if (fresh) {
                    have = true;
                    if (!init) finish();   // failed initial state, or nothing to integrate
                }
            }
        }
        if (!__any(have)) break;
        // ---- per-row checks before the attempt (the rows that stop here write their mode)
        if (have) {
            if (++iters > a.maxiters) ret = SBZ_RET_MAXITERS;
            else if (t + dt >= t1 - 1e-12 * fabs(t1)) dt = t1 - t;
            else if (dt < 1e-14 * fabs(t)) ret = SBZ_RET_DTLESSTHANMIN;
            if (ret != SBZ_RET_SUCCESS) finish();
        }
        const bool live = have;
        const double h = live ? dt : 1e-3 * (t1 - kt[0]);   // rows without a mode: a harmless step
        double* const Un = L.v[ub];
        double* const U1 = L.v[ub ^ 1];
        double* const F0 = L.v[fb];
        double* const F1 = L.v[fb ^ 1];
        const double ig = 1.0 / (SBZ_R5_GAMMA * h);
        const double idt = 1.0 / h;
        // background at the six distinct stage times, lanes 0..5 of every row
        if (c.lane < 6) {
            const double cl = (c.lane == 0) ? 0.0 : (c.lane == 1) ? SBZ_R5_C2 : (c.lane == 2) ? SBZ_R5_C3
                            : (c.lane == 3) ? SBZ_R5_C4 : (c.lane == 4) ? SBZ_R5_C5 : 1.0;
            double Bq[SBZ_NB], dBq[SBZ_NB];
            sbz_bg_dual(P, kt, nk, bidx, t + cl * h, Bq, dBq);
#pragma unroll
            for (int q = 0; q < SBZ_NB; ++q) L.B[c.lane][q] = Bq[q];
            if (c.lane == 0) {
#pragma unroll
                for (int q = 0; q < SBZ_NB; ++q) L.dB[q] = dBq[q];
            }
        }
        __syncthreads();
        const double* B0 = L.B[0];
        if (c.isC) {
            const SbzChainRow r = sbz_chain_row(c, B0, L.Cs, k);
            csk = r.sk;
            double xcs = 0.0;
#define SBZ_CG_FACTOR(lo, hi, len) \
            if (SBZ_IN_GROUP(lo, hi)) xcs = sbz_chain_factor<len>(r, ig, cinv);
            SBZ_CHAIN_GROUPS(SBZ_CG_FACTOR)
#undef SBZ_CG_FACTOR
            L.C[c.lane] = xcs;
        }
        __syncthreads();
        const double xc = (c.isB && c.bchain >= 0) ? L.C[c.bchain] : 0.0;
        double al, be;
        sbz_wb_w(c.lane, B0, L.Cs, k, ig, xc, L.M, row, cpl, al, be, wb);
        bool ok = sbz_wb_lu<BC>(c.lane, row, rpo);
        sbz_wb_unpack(c.lane, row, lrow, urow);
        if constexpr (!OWNSCALE) sbz_wb_rpb<BC>(rpo, rpb);
        sbz_wb_tri2_r<BC, OWNSCALE>(lrow, urow, rpb, rpo, al, be);
        wb.z1 = c.isB ? al : 0.0;
        wb.z2 = c.isB ? be : 0.0;
        {
            const double c11 = 1.0 - RO::sum(wb.p * wb.z1), c12 = -RO::sum(wb.p * wb.z2);
            const double c21 = -RO::sum(wb.q * wb.z1), c22 = 1.0 - RO::sum(wb.q * wb.z2);
            const double det = c11 * c22 - c12 * c21;
            ok = ok && (det != 0.0) && (det == det) && (fabs(det) < 1e300);
            const double idet = 1.0 / det;
            wb.c11 = c22 * idet;
            wb.c12 = -c12 * idet;
            wb.c21 = -c21 * idet;
            wb.c22 = c11 * idet;
        }
        ok = RO::all(ok || !c.isB, g);
        sbz_wave_apply<LOADS_FIRST>(c, B0, L.dB, L.Cs, k, Un, DT);  // ∂f/∂τ at (t, u_n)
        if (live) st_j++;
        __syncthreads();
        double EEst = __builtin_inf();
        // stages 1..8, then pass 9 evaluates f(t + h, u_{n+1}) for the next step
        for (int s = 1; s <= 9; ++s) {
            double* Ks = (s <= 8) ? L.v[5 + s] : F1;
            if (s == 1) {
                if constexpr (LOADS_FIRST) {   // every load first, masked stores
                    double f0[OWN], d0[OWN];
#pragma unroll
                    for (int r = 0; r < OWN; ++r) {
                        const int i = (c.lane + LN * r < SBZ_N) ? c.lane + LN * r : SBZ_N - 1;
                        f0[r] = F0[i];
                        d0[r] = DT[i];
                    }
#pragma unroll
                    for (int r = 0; r < OWN; ++r)
                        if (c.lane + LN * r < SBZ_N) Ks[c.lane + LN * r] = f0[r] + h * SBZ_R5_D1 * d0[r];
                } else {
                    for (int i = c.lane; i < SBZ_N; i += LN) Ks[i] = F0[i] + h * SBZ_R5_D1 * DT[i];
                }
            } else {
                double* in = (s <= 8) ? US : U1;
                double ck[OWN];
                if (s <= 8) {
                    if constexpr (FUSED) SBZ_STAGE_SWITCH_LN(sbz_stage_in_f, LN, c.lane, L.v, Un, US, idt, ck)
                    else SBZ_STAGE_SWITCH_LN(sbz_stage_in, LN, c.lane, L.v, Un, US)
                }
                __syncthreads();
                sbz_wave_apply<LOADS_FIRST>(c, L.B[sbz_stage_bg[s]], nullptr, L.Cs, k, in, Ks);
                // pass 9 runs for a row that rejected when another row accepted: not counted, so
                // a mode's statistics do not depend on the modes that share its wavefront
                if (live && (s <= 8 || EEst <= 1.0)) st_f++;
                __syncthreads();
                if (s == 9) break;
                if constexpr (FUSED) SBZ_STAGE_SWITCH_LN(sbz_stage_rhs_f, LN, c.lane, h, DT, Ks, ck)
                else SBZ_STAGE_SWITCH_LN(sbz_stage_rhs, LN, c.lane, L.v, h, idt, DT, Ks)
            }
            __syncthreads();
            // solve W x = Ks: chain forward sweeps, border solve, chain back-substitution
            if (c.isC) {
                double y = 0.0;
#define SBZ_CG_FWD(lo, hi, len) \
                if (SBZ_IN_GROUP(lo, hi)) y = sbz_chain_fwd<len>(csk, cinv, Ks + c.cbase);
                SBZ_CHAIN_GROUPS(SBZ_CG_FWD)
#undef SBZ_CG_FWD
                L.C[c.lane] = y;
            }
            __syncthreads();
            {
                double b = c.isB ? Ks[c.bcomp] : 0.0;
                if (c.isB && c.bchain >= 0) b -= cpl * L.C[c.bchain];
                sbz_wb_tri_r<BC, OWNSCALE>(lrow, urow, rpb, rpo, b);   // y = A^-1 b
                const double s1 = RO::sum(wb.p * b), s2 = RO::sum(wb.q * b);
                const double x = b + wb.z1 * (wb.c11 * s1 + wb.c12 * s2) + wb.z2 * (wb.c21 * s1 + wb.c22 * s2);
                if (c.isB) {
                    Ks[c.bcomp] = x;
                    L.X[c.lane] = x;
                }
            }
            __syncthreads();
            if (c.isC) {
                const double xatt = L.X[c.catt];
#define SBZ_CG_BWD(lo, hi, len) \
                if (SBZ_IN_GROUP(lo, hi)) sbz_chain_bwd<len>(csk, cinv, xatt, Ks + c.cbase);
                SBZ_CHAIN_GROUPS(SBZ_CG_BWD)
#undef SBZ_CG_BWD
            }
            __syncthreads();
            if (s == 8) {
                // u_{n+1} = stage-8 input + k8; error estimate k8
                double ss = 0.0;
                bool fin = true;
                if constexpr (LOADS_FIRST) {   // every load first, as for stage 1
                    double us[OWN], kv[OWN], uo[OWN];
#pragma unroll
                    for (int r = 0; r < OWN; ++r) {
                        const int i = (c.lane + LN * r < SBZ_N) ? c.lane + LN * r : SBZ_N - 1;
                        us[r] = US[i];
                        kv[r] = Ks[i];
                        uo[r] = Un[i];
                    }
#pragma unroll
                    for (int r = 0; r < OWN; ++r) {
                        const int i = c.lane + LN * r;
                        if (i < SBZ_N) {
                            const double un1 = us[r] + kv[r];
                            U1[i] = un1;
                            const double sc = a.atol + fmax(fabs(uo[r]), fabs(un1)) * a.rtol;
                            const double q = kv[r] / sc;
                            ss += q * q;
                            if (!(un1 == un1) || fabs(un1) > 1e300) fin = false;
                        }
                    }
                } else {
                    for (int i = c.lane; i < SBZ_N; i += LN) {
                        const double un1 = US[i] + Ks[i];
                        U1[i] = un1;
                        const double sc = a.atol + fmax(fabs(Un[i]), fabs(un1)) * a.rtol;
                        const double q = Ks[i] / sc;
                        ss += q * q;
                        if (!(un1 == un1) || fabs(un1) > 1e300) fin = false;
                    }
                }
                ss = RO::sum(ss);
                const bool fin_row = RO::all(fin, g);   // a ballot: evaluated by every lane
                EEst = (ok && fin_row) ? sqrt(ss / SBZ_N) : __builtin_inf();
                if (!__any(live && EEst <= 1.0)) break;   // no row accepts: no f(t + h) needed
            }
        }
        __syncthreads();
        // the controller's two powers EEst^β1 and qold^β2, on lanes 0 and 1 of every row
        const double pw = pow((c.lane == 1) ? qold : ((EEst < __builtin_inf()) ? EEst : 1.0),
                              (c.lane == 1) ? SBZ_CTL_BETA2 : SBZ_CTL_BETA1);
        const double q11 = BC::template at<0>(pw), qob = BC::template at<1>(pw);
        if (live) {
            if (!(EEst <= 1.0)) {
                const double q11r = (EEst < __builtin_inf()) ? q11 : 1.0 / SBZ_CTL_QMIN;
                dt /= fmin(1.0 / SBZ_CTL_QMIN, q11r / SBZ_CTL_GAMMA);
                st_rej++;
            } else {
                const double q = (EEst == 0.0) ? 1.0 / SBZ_CTL_QMAX
                                               : fmax(1.0 / SBZ_CTL_QMAX,
                                                      fmin(1.0 / SBZ_CTL_QMIN, q11 / qob / SBZ_CTL_GAMMA));
                qold = fmax(EEst, SBZ_CTL_QOLDINIT);
                double tn = t + dt;
                if (tn >= t1 - 1e-14 * fabs(t1)) tn = t1;
                t = tn;
                ub ^= 1;
                fb ^= 1;
                while (kidx < nk - 2 && t >= kt[kidx + 1]) ++kidx;
                st_acc++;
                dt = dt / q;
                if (!(t < t1)) finish();
            }
        }
        __syncthreads();
    }
}

template <bool FUSED, int LN>
SBZ_DEV void sbz_waverow_body(const SbzSolveArgs& a) {
    __shared__ WaveRow R4[64 / LN];
    sbz_waverow_run<FUSED, LN>(a, R4);
}
#endif

// ============================================================================================
// Tangent solve (SBZ_ALGO_TANGENT): u and s = ∂u/∂log θ integrated together, one mode per
// wavefront.  The augmented system  u' = J u,  s' = J s + J_θ u  is linear with the block
// lower-triangular Jacobian [[J, 0], [J_θ, J]], so every Rosenbrock stage needs W x = b twice
// with the one factorization of W = I/(γh) - J:  W x_u = b_u,  W x_s = b_s + J_θ x_u.  J_θ v is
// (f(v; B₊) - f(v; B₋)) / (2h), with B± the background at the same τ from the parameter sets of
// log θ ± h (solved tightly by the host), and ∂(J_θ u)/∂τ is the same difference of ∂f/∂τ.  The
// step sequence is the primal's: the error norm and the controller read u only, so u equals the
// plain solve's and s is the derivative of the discrete solution along that step sequence.
// States only: no saves or sources.
// ============================================================================================
#if SBZ_TAN_OK
struct WaveTanLDS {
    double v[15][SBZ_N];           // primal: 0,1 U; 2,3 F; 4 DT; 5 US; 6..13 K1..K8; 14 scratch
    double w[15][SBZ_N];           // tangent, same slots
    double B[3][7][SBZ_NB];        // background of the sets θ, θ+, θ- at the stage times (slot 6: spare)
    double dB[3][SBZ_NB];          // their τ-derivatives at t_n
    double Cs[SBZ_NC];
    double M0[SBZ_NBORD * SBZ_WSTRIDE];   // border staging
    double LU[(1 + SBZ_LU_SPLIT) * SBZ_NBORD * SBZ_WSTRIDE];   // factored border block (sbz_wb_store)
    double X[SBZ_NBORD];
    double C[SBZ_NCHAIN];
    int mode;
};

SBZ_DEV void sbz_wave_tan_body(const SbzSolveArgs& a) {
    __shared__ WaveTanLDS L;
    WaveCtx c;
    c.lane = threadIdx.x;
    c.isB = c.lane < SBZ_NBORD;
    c.bcomp = c.isB ? sbz_bord_comp[c.lane] : 0;
    c.bchain = c.isB ? sbz_bord_chain[c.lane] : -1;
    c.isC = c.lane < SBZ_NCHAIN;
    const int ci = c.isC ? c.lane : 0;
    c.ckind = sbz_chain_kind[ci];
    c.cL = sbz_chain_L[ci];
    c.cbin = sbz_chain_bin[ci];
    c.catt = sbz_chain_att[ci];
    c.cattc = sbz_chain_attc[ci];
    c.cbase = sbz_chain_base[ci];
    double* const DT = L.v[4];
    double* const US = L.v[5];
    double* const TMP = L.v[14];
    double* const DTs = L.w[4];
    double* const SS = L.w[5];
    double* const TMPs = L.w[14];
    __syncthreads();
    for (;;) {
        if (c.lane == 0) L.mode = sbz_pull(a, (int)(blockIdx.x % (unsigned)a.n_xcd));
        __syncthreads();
        const int qi = L.mode;
        __syncthreads();
        if (qi < 0) break;
        const int inst = a.order ? a.order[qi] : qi;
        const int set0 = a.inst_set[inst], setp = a.inst_setp[inst], setm = a.inst_setm[inst];
        const double tsc = a.inst_tscale[inst];   // 1 / (2h)
        const double* P = a.params + (size_t)set0 * SBZ_NPAR;
        const double* kt = a.knots + a.knot_off[set0];
        const int nk = a.knot_n[set0];
        // the table this lane evaluates at the stage times: lanes 0..5 θ, 6..11 θ + h, 12..17 θ - h
        const int bs = (c.lane < 18) ? c.lane / 6 : 0;
        const int bset = (bs == 0) ? set0 : (bs == 1) ? setp : setm;
        const double* Pb = a.params + (size_t)bset * SBZ_NPAR;
        const double* ktb = a.knots + a.knot_off[bset];
        const int nkb = a.knot_n[bset];
        const double k = a.inst_k[inst];
        const double t1 = kt[nk - 1];
        double t = a.inst_t0[inst];
        int kidx = sbz_find_knot(kt, nk, t);
        int bidx = sbz_find_knot(ktb, nkb, t);
        int ub = 0, fb = 2;
        long long st_acc = 0, st_rej = 0, st_f = 0, st_j = 0;
        int ret = SBZ_RET_SUCCESS;
        // the three backgrounds at the start time (slot 6), the constants
        if (c.lane < 3) {
            const int sq = (c.lane == 0) ? set0 : (c.lane == 1) ? setp : setm;
            const double* Pq = a.params + (size_t)sq * SBZ_NPAR;
            const double* ktq = a.knots + a.knot_off[sq];
            const int nkq = a.knot_n[sq];
            double Bx[SBZ_NB];
            int kk = sbz_find_knot(ktq, nkq, t);
            sbz_bg_plain(Pq, ktq, nkq, kk, t, Bx);
#pragma unroll
            for (int q = 0; q < SBZ_NB; ++q) L.B[c.lane][6][q] = Bx[q];
            if (c.lane == 0) {
                double Cs[SBZ_NC];
                sbz_consts(P, Cs);
                sbz_consts_k(Cs, k);
#pragma unroll
                for (int q = 0; q < SBZ_NC; ++q) L.Cs[q] = Cs[q];
            }
        }
        __syncthreads();
        // initial state u0 and its tangent s0 = (u0(θ + h) - u0(θ - h)) / (2h)
        if (c.lane == 0) {
            sbz_ic(L.B[0][6], L.Cs, k, L.v[0]);
            sbz_ic(L.B[1][6], L.Cs, k, L.w[6]);
            sbz_ic(L.B[2][6], L.Cs, k, L.w[7]);
        }
        __syncthreads();
        {
            bool fin = (k > 0.0) && (k == k);
            for (int i = c.lane; i < SBZ_N; i += 64) {
                L.w[0][i] = (L.w[6][i] - L.w[7][i]) * tsc;
                const double v = L.v[0][i], ws = L.w[0][i];
                if (!(v == v) || fabs(v) > 1e300 || !(ws == ws) || fabs(ws) > 1e300) fin = false;
            }
            if (!__all(fin)) ret = SBZ_RET_INITIALFAILURE;
        }
        for (int i = c.lane; i < SBZ_N; i += 64)
#pragma unroll
            for (int j = 6; j <= 13; ++j) {
                L.v[j][i] = 0.0;
                L.w[j][i] = 0.0;
            }
        __syncthreads();
        // g(S, U) = f(S; B) + (f(U; B+) - f(U; B-)) / (2h), the tangent block's right-hand side
        auto tangent_rhs = [&](const double* B0q, const double* Bpq, const double* Bmq, const double* Sin,
                               const double* Uin, double* out) {
            sbz_wave_apply(c, B0q, nullptr, L.Cs, k, Sin, out);
            sbz_wave_apply(c, Bpq, nullptr, L.Cs, k, Uin, TMP);
            sbz_wave_apply(c, Bmq, nullptr, L.Cs, k, Uin, TMPs);
            __syncthreads();
            for (int i = c.lane; i < SBZ_N; i += 64) out[i] += (TMP[i] - TMPs[i]) * tsc;
            __syncthreads();
        };
        // f and g at the start: F0 = v[2], G0 = w[2]
        sbz_wave_apply(c, L.B[0][6], nullptr, L.Cs, k, L.v[0], L.v[2]);
        tangent_rhs(L.B[0][6], L.B[1][6], L.B[2][6], L.w[0], L.v[0], L.w[2]);
        double dt = 0.0, qold = SBZ_CTL_QOLDINIT;
        if (ret == SBZ_RET_SUCCESS && t1 > t) {
            // Hairer–Wanner initial step from the primal, as the plain solve
            double d0 = 0.0, d1 = 0.0;
            for (int i = c.lane; i < SBZ_N; i += 64) {
                const double skv = a.atol + fabs(L.v[0][i]) * a.rtol;
                d0 += (L.v[0][i] / skv) * (L.v[0][i] / skv);
                d1 += (L.v[2][i] / skv) * (L.v[2][i] / skv);
            }
            d0 = sqrt(sbz_wave_sum(d0) / SBZ_N);
            d1 = sqrt(sbz_wave_sum(d1) / SBZ_N);
            double dt0 = (d0 < 1e-5 || d1 < 1e-5) ? 1e-6 : 0.01 * d0 / d1;
            dt0 = fmin(dt0, t1 - t);
            for (int i = c.lane; i < SBZ_N; i += 64) L.v[1][i] = L.v[0][i] + dt0 * L.v[2][i];
            if (c.lane == 0) {
                double Bx[SBZ_NB];
                int kk = kidx;
                sbz_bg_plain(P, kt, nk, kk, t + dt0, Bx);
#pragma unroll
                for (int q = 0; q < SBZ_NB; ++q) L.B[0][6][q] = Bx[q];
            }
            __syncthreads();
            sbz_wave_apply(c, L.B[0][6], nullptr, L.Cs, k, L.v[1], L.v[3]);
            __syncthreads();
            double d2 = 0.0;
            for (int i = c.lane; i < SBZ_N; i += 64) {
                const double skv = a.atol + fabs(L.v[0][i]) * a.rtol;
                const double r = (L.v[3][i] - L.v[2][i]) / skv;
                d2 += r * r;
            }
            d2 = sqrt(sbz_wave_sum(d2) / SBZ_N) / dt0;
            const double mx = fmax(d1, d2);
            const double dt1 = (mx <= 1e-15) ? fmax(1e-6, dt0 * 1e-3) : pow(0.01 / mx, 1.0 / 6.0);
            dt = fmin(fmin(100.0 * dt0, dt1), t1 - t);
            st_f += 2;
        }
        double row[SBZ_NBORD];
        double cpl = 0.0;
        SbzWb wb;
        double cinv[SBZ_MCHAIN];
        double csk = 0.0;
        long long iters = 0;
        // W x = b in place on the LDS vector b: chain forward sweeps, border solve, back-substitution
        auto solve_W = [&](double* bv) {
            if (c.isC) {
                double y = 0.0;
#define SBZ_CG_FWD(lo, hi, len) \
                if (SBZ_IN_GROUP(lo, hi)) y = sbz_chain_fwd<len>(csk, cinv, bv + c.cbase);
                SBZ_CHAIN_GROUPS(SBZ_CG_FWD)
#undef SBZ_CG_FWD
                L.C[c.lane] = y;
            }
            __syncthreads();
            {
                double b = c.isB ? bv[c.bcomp] : 0.0;
                if (c.isB && c.bchain >= 0) b -= cpl * L.C[c.bchain];
                sbz_wb_tri<SBZ_LU_SPLIT, SbzBcWave>(c.lane, L.LU, b);
                const double s1 = sbz_wave_sum(wb.p * b), s2 = sbz_wave_sum(wb.q * b);
                const double x = b + wb.z1 * (wb.c11 * s1 + wb.c12 * s2) + wb.z2 * (wb.c21 * s1 + wb.c22 * s2);
                if (c.isB) {
                    bv[c.bcomp] = x;
                    L.X[c.lane] = x;
                }
            }
            __syncthreads();
            if (c.isC) {
                const double xatt = L.X[c.catt];
#define SBZ_CG_BWD(lo, hi, len) \
                if (SBZ_IN_GROUP(lo, hi)) sbz_chain_bwd<len>(csk, cinv, xatt, bv + c.cbase);
                SBZ_CHAIN_GROUPS(SBZ_CG_BWD)
#undef SBZ_CG_BWD
            }
            __syncthreads();
        };
        for (;;) {
            if (!(ret == SBZ_RET_SUCCESS && t < t1)) break;
            double* const Un = L.v[ub];
            double* const U1 = L.v[ub ^ 1];
            double* const F0 = L.v[fb];
            double* const F1 = L.v[fb ^ 1];
            double* const Sn = L.w[ub];
            double* const S1 = L.w[ub ^ 1];
            double* const G0 = L.w[fb];
            double* const G1 = L.w[fb ^ 1];
            if (++iters > a.maxiters) { ret = SBZ_RET_MAXITERS; break; }
            if (t + dt >= t1 - 1e-12 * fabs(t1)) dt = t1 - t;
            else if (dt < 1e-14 * fabs(t)) { ret = SBZ_RET_DTLESSTHANMIN; break; }
            const double ig = 1.0 / (SBZ_R5_GAMMA * dt);
            const double idt = 1.0 / dt;
            // the three backgrounds at the six stage times, lanes 0..17 (dual: d/dτ at t_n)
            if (c.lane < 18) {
                const int q = c.lane - 6 * bs;
                const double cl = (q == 0) ? 0.0 : (q == 1) ? SBZ_R5_C2 : (q == 2) ? SBZ_R5_C3
                                : (q == 3) ? SBZ_R5_C4 : (q == 4) ? SBZ_R5_C5 : 1.0;
                double Bq[SBZ_NB], dBq[SBZ_NB];
                sbz_bg_dual(Pb, ktb, nkb, bidx, t + cl * dt, Bq, dBq);
#pragma unroll
                for (int r = 0; r < SBZ_NB; ++r) L.B[bs][q][r] = Bq[r];
                if (q == 0) {
#pragma unroll
                    for (int r = 0; r < SBZ_NB; ++r) L.dB[bs][r] = dBq[r];
                }
            }
            __syncthreads();
            const double* B0 = L.B[0][0];
            if (c.isC) {
                const SbzChainRow r = sbz_chain_row(c, B0, L.Cs, k);
                csk = r.sk;
                double xcs = 0.0;
#define SBZ_CG_FACTOR(lo, hi, len) \
                if (SBZ_IN_GROUP(lo, hi)) xcs = sbz_chain_factor<len>(r, ig, cinv);
                SBZ_CHAIN_GROUPS(SBZ_CG_FACTOR)
#undef SBZ_CG_FACTOR
                L.C[c.lane] = xcs;
            }
            __syncthreads();
            const double xc = (c.isB && c.bchain >= 0) ? L.C[c.bchain] : 0.0;
            double al, be;
            sbz_wb_w(c.lane, B0, L.Cs, k, ig, xc, L.M0, row, cpl, al, be, wb);
            double rpo;
            bool ok = sbz_wb_lu<SbzBcWave>(c.lane, row, rpo);
            sbz_wb_store<SBZ_LU_SPLIT>(c.lane, row, rpo, L.LU);
            sbz_wb_tri2<SBZ_LU_SPLIT, SbzBcWave>(c.lane, L.LU, al, be);
            wb.z1 = c.isB ? al : 0.0;
            wb.z2 = c.isB ? be : 0.0;
            {
                const double c11 = 1.0 - sbz_wave_sum(wb.p * wb.z1), c12 = -sbz_wave_sum(wb.p * wb.z2);
                const double c21 = -sbz_wave_sum(wb.q * wb.z1), c22 = 1.0 - sbz_wave_sum(wb.q * wb.z2);
                const double det = c11 * c22 - c12 * c21;
                ok = ok && (det != 0.0) && (det == det) && (fabs(det) < 1e300);
                const double idet = 1.0 / det;
                wb.c11 = c22 * idet;
                wb.c12 = -c12 * idet;
                wb.c21 = -c21 * idet;
                wb.c22 = c11 * idet;
            }
            ok = __all(ok || !c.isB);
            // ∂f/∂τ of the primal, and of the tangent block: ∂f/∂τ(s) + (∂f/∂τ(u; θ+) - ∂f/∂τ(u; θ-)) / (2h)
            sbz_wave_apply(c, B0, L.dB[0], L.Cs, k, Un, DT);
            sbz_wave_apply(c, B0, L.dB[0], L.Cs, k, Sn, DTs);
            sbz_wave_apply(c, L.B[1][0], L.dB[1], L.Cs, k, Un, TMP);
            sbz_wave_apply(c, L.B[2][0], L.dB[2], L.Cs, k, Un, TMPs);
            __syncthreads();
            for (int i = c.lane; i < SBZ_N; i += 64) DTs[i] += (TMP[i] - TMPs[i]) * tsc;
            st_j++;
            __syncthreads();
            double EEst = __builtin_inf();
            if (ok) {
                for (int s = 1; s <= 9; ++s) {
                    double* Ks = (s <= 8) ? L.v[5 + s] : F1;
                    double* KSs = (s <= 8) ? L.w[5 + s] : G1;
                    if (s == 1) {
                        for (int i = c.lane; i < SBZ_N; i += 64) {
                            Ks[i] = F0[i] + dt * SBZ_R5_D1 * DT[i];
                            KSs[i] = G0[i] + dt * SBZ_R5_D1 * DTs[i];
                        }
                    } else {
                        double* inU = (s <= 8) ? US : U1;
                        double* inS = (s <= 8) ? SS : S1;
                        if (s <= 8) {
                            SBZ_STAGE_SWITCH_LN(sbz_stage_in, 64, c.lane, L.v, Un, US)
                            SBZ_STAGE_SWITCH_LN(sbz_stage_in, 64, c.lane, L.w, Sn, SS)
                        }
                        __syncthreads();
                        const int bq = sbz_stage_bg[s];
                        sbz_wave_apply(c, L.B[0][bq], nullptr, L.Cs, k, inU, Ks);
                        tangent_rhs(L.B[0][bq], L.B[1][bq], L.B[2][bq], inS, inU, KSs);
                        st_f++;
                        if (s == 9) break;
                        SBZ_STAGE_SWITCH_LN(sbz_stage_rhs, 64, c.lane, L.v, dt, idt, DT, Ks)
                        SBZ_STAGE_SWITCH_LN(sbz_stage_rhs, 64, c.lane, L.w, dt, idt, DTs, KSs)
                    }
                    __syncthreads();
                    solve_W(Ks);   // W x_u = b_u
                    // W x_s = b_s + J_θ x_u, J_θ at t_n
                    sbz_wave_apply(c, L.B[1][0], nullptr, L.Cs, k, Ks, TMP);
                    sbz_wave_apply(c, L.B[2][0], nullptr, L.Cs, k, Ks, TMPs);
                    __syncthreads();
                    for (int i = c.lane; i < SBZ_N; i += 64) KSs[i] += (TMP[i] - TMPs[i]) * tsc;
                    __syncthreads();
                    solve_W(KSs);
                    if (s == 8) {
                        // u_{n+1}, s_{n+1}; the error estimate of the primal only
                        double ss = 0.0;
                        bool fin = true;
                        for (int i = c.lane; i < SBZ_N; i += 64) {
                            const double un1 = US[i] + Ks[i], sn1 = SS[i] + KSs[i];
                            U1[i] = un1;
                            S1[i] = sn1;
                            const double scl = a.atol + fmax(fabs(Un[i]), fabs(un1)) * a.rtol;
                            const double r = Ks[i] / scl;
                            ss += r * r;
                            if (!(un1 == un1) || fabs(un1) > 1e300 || !(sn1 == sn1) || fabs(sn1) > 1e300) fin = false;
                        }
                        ss = sbz_wave_sum(ss);
                        EEst = __all(fin) ? sqrt(ss / SBZ_N) : __builtin_inf();
                        if (!(EEst <= 1.0)) break;
                    }
                }
            }
            __syncthreads();
            const double q11 = (EEst < __builtin_inf()) ? pow(EEst, SBZ_CTL_BETA1) : 1.0;
            if (!(EEst <= 1.0)) {
                const double q11r = (EEst < __builtin_inf()) ? q11 : 1.0 / SBZ_CTL_QMIN;
                dt /= fmin(1.0 / SBZ_CTL_QMIN, q11r / SBZ_CTL_GAMMA);
                st_rej++;
                if (!(EEst < __builtin_inf())) {
                    for (int i = c.lane; i < SBZ_N; i += 64)
#pragma unroll
                        for (int j = 6; j <= 13; ++j) {
                            L.v[j][i] = 0.0;
                            L.w[j][i] = 0.0;
                        }
                    __syncthreads();
                }
                continue;
            }
            const double q = (EEst == 0.0) ? 1.0 / SBZ_CTL_QMAX
                                           : fmax(1.0 / SBZ_CTL_QMAX,
                                                  fmin(1.0 / SBZ_CTL_QMIN, q11 / pow(qold, SBZ_CTL_BETA2) / SBZ_CTL_GAMMA));
            qold = fmax(EEst, SBZ_CTL_QOLDINIT);
            double tn = t + dt;
            if (tn >= t1 - 1e-14 * fabs(t1)) tn = t1;
            t = tn;
            ub ^= 1;
            fb ^= 1;
            while (kidx < nk - 2 && t >= kt[kidx + 1]) ++kidx;
            st_acc++;
            dt = dt / q;
        }
        __syncthreads();
        {
            double* ou = a.out_u + (size_t)inst * SBZ_N;
            double* os = a.out_s + (size_t)inst * SBZ_N;
            double* of = a.out_fend + (size_t)inst * SBZ_N;
            for (int i = c.lane; i < SBZ_N; i += 64) {
                ou[i] = L.v[ub][i];
                os[i] = L.w[ub][i];
                of[i] = L.v[fb][i];   // f(t, u) at the end: the last pass 9, or the start's f
            }
        }
        if (c.lane == 0) {
            a.retcode[inst] = ret;
            long long* st = a.stats + (size_t)inst * 6;
            st[0] = st_acc;
            st[1] = st_rej;
            st[2] = st_f;
            st[3] = st_j;
            st[4] = st_j;
            st[5] = 16 * st_j;
        }
        __syncthreads();
    }
}
#endif

// The kernels of a single-model code object (a mixed-model code object compiles each model's
// device code in its own namespace and defines its own kernel, sbz_multi_kernel).
#ifndef SBZ_MULTI_FAMILY
// The same integrator compiled twice: for one resident wavefront per SIMD (all 512 registers
// per lane, fewest spills: lowest latency per mode) and for two (half the registers, a few more
// spills: the second wavefront hides the first's latency when there are more instances than
// SIMDs).  SBZ_ALGO_WAVE / SBZ_ALGO_WAVE2 select them.
extern "C" __global__ void __launch_bounds__(64) sbz_wave_kernel(SbzSolveArgs a) { sbz_wave_body<true>(a); }
extern "C" __global__ void __launch_bounds__(64, 2) sbz_wave_kernel_o2(SbzSolveArgs a) { sbz_wave_body<false>(a); }
// The one-wavefront kernel with the other border-factor layout (split where the model's default is
// combined and vice versa): fewer instructions per solve against less LDS (SBZ_ALGO_WAVE_ALTLU).
extern "C" __global__ void __launch_bounds__(64) sbz_wave_kernel_altlu(SbzSolveArgs a) {
    sbz_wave_body<true, !SBZ_LU_SPLIT>(a);
}

#if SBZ_TAN_OK
// Tangent solve, one mode and its parameter derivative per wavefront.
extern "C" __global__ void __launch_bounds__(64) sbz_wave_tan_kernel(SbzSolveArgs a) { sbz_wave_tan_body(a); }
#endif

#if SBZ_WAVE4_OK
// Four modes per wavefront, one resident wavefront per SIMD (all 512 registers per lane).
extern "C" __global__ void __launch_bounds__(64) sbz_wave4_kernel(SbzSolveArgs a) { sbz_waverow_body<true, 16>(a); }
#endif
#if SBZ_WAVE2H_OK
// Two modes per wavefront, for models whose border or chains exceed 16 lanes but fit 32.
extern "C" __global__ void __launch_bounds__(64) sbz_wave2h_kernel(SbzSolveArgs a) { sbz_waverow_body<true, 32>(a); }
#endif

// Source background table for the wave kernels: one thread per (set, save time), rows as
// sbz_srcbg_row.  Launched before the solve when the plan saves CMB sources.
extern "C" __global__ void __launch_bounds__(64) sbz_srcbg_kernel(const double* __restrict__ params,
                                                                  const double* __restrict__ knots,
                                                                  const long long* __restrict__ knot_off,
                                                                  const int* __restrict__ knot_n,
                                                                  const double* __restrict__ saveat, int saveat_stride,
                                                                  int n_save, long long n_rows, double* __restrict__ out) {
    const long long g = (long long)blockIdx.x * 64 + threadIdx.x;
    if (g >= n_rows) return;
    // a set index below n_sets that never received a background table has no knots: its rows are
    // never read (instances refer to valid sets only) and are skipped instead of reading past them
    if (knot_n[g / n_save] < 2) return;
    sbz_srcbg_row(params, knots, knot_off, knot_n, saveat, saveat_stride, n_save, g, out);
}

// Plan outputs start as NaN at every run: a save time that a mode never reached (it failed first)
// then reads NaN instead of stale memory.  A kernel rather than a memset, so that a run can be
// captured in a HIP graph.
extern "C" __global__ void __launch_bounds__(256) sbz_fill_kernel(double* __restrict__ p, long long n, double v) {
    for (long long i = (long long)blockIdx.x * 256 + threadIdx.x; i < n; i += (long long)gridDim.x * 256) p[i] = v;
}

// Zeroes the work-queue counter before a solve launch (a kernel node, not a memset node, in graphs).
extern "C" __global__ void sbz_queue_reset_kernel(unsigned int* __restrict__ q, int n) {
    if (blockIdx.x == 0 && (int)threadIdx.x < n) q[threadIdx.x] = 0u;
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
    double Cs[SBZ_NC], B[SBZ_NB];
    sbz_consts(P, Cs);
    int idx = nk - 2;
    sbz_bg_plain(P, kt, nk, idx, kt[nk - 1], B);
    CSV u{u_final + (size_t)i * SBZ_N, 1};
    double dm = sbz_dm(B, Cs, k, u);
    double p0 = 2.0 * sbzc::PI * sbzc::PI / (k * k * k) * P[SBZ_P_AS] * pow(k / P[SBZ_P_KPIVOT], P[SBZ_P_NS] - 1.0);
    pk[i] = p0 * dm * dm;
}

// P(k, τ0) and d log P / d log θ of tangent solves (sbz_wave_tan_kernel's u, s, f(τ0)).  τ0 itself
// depends on θ (a(τ0) = 1), so the total derivative of the final state is s + f(τ0) dτ0/dlog θ,
// with dτ0/dlog θ from the τ0 of the sets θ ± h; and Δ_m (linear in u) reads the background at τ0
// of each set:  dΔ_m/dlog θ = (Δ_m(u + h ṡ; B₊(τ0₊)) - Δ_m(u - h ṡ; B₋(τ0₋))) / (2h), ṡ the total
// tangent.  d log P = 2 dΔ_m / Δ_m (P0 does not depend on θ for the parameters this serves).
extern "C" __global__ void sbz_pk_tan_kernel(const double* __restrict__ params, const double* __restrict__ knots,
                                             const long long* __restrict__ knot_off, const int* __restrict__ knot_n,
                                             const int* __restrict__ inst_set, const int* __restrict__ inst_setp,
                                             const int* __restrict__ inst_setm, const double* __restrict__ inst_tscale,
                                             const double* __restrict__ inst_k, const double* __restrict__ u_final,
                                             const double* __restrict__ s_final, const double* __restrict__ f_end,
                                             int n_inst, double* __restrict__ pk, double* __restrict__ dlogpk) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n_inst) return;
    const int sets[3] = {inst_set[i], inst_setp[i], inst_setm[i]};
    const double k = inst_k[i], tsc = inst_tscale[i], h = 0.5 / tsc;
    const double* P = params + (size_t)sets[0] * SBZ_NPAR;
    double Cs[SBZ_NC];
    sbz_consts(P, Cs);
    CSV u{u_final + (size_t)i * SBZ_N, 1}, sv{s_final + (size_t)i * SBZ_N, 1}, fv{f_end + (size_t)i * SBZ_N, 1};
    double dm[3][3], tau0[3];   // [set][u, s, f]
    for (int q = 0; q < 3; ++q) {
        const double* Pq = params + (size_t)sets[q] * SBZ_NPAR;
        const double* kt = knots + knot_off[sets[q]];
        const int nk = knot_n[sets[q]];
        tau0[q] = kt[nk - 1];
        double B[SBZ_NB];
        int idx = nk - 2;
        sbz_bg_plain(Pq, kt, nk, idx, kt[nk - 1], B);
        dm[q][0] = sbz_dm(B, Cs, k, u);
        dm[q][1] = sbz_dm(B, Cs, k, sv);
        dm[q][2] = sbz_dm(B, Cs, k, fv);
    }
    const double dtau0 = (tau0[1] - tau0[2]) * tsc;
    const double Dp = dm[1][0] + h * (dm[1][1] + dtau0 * dm[1][2]);
    const double Dm = dm[2][0] - h * (dm[2][1] + dtau0 * dm[2][2]);
    const double D0 = dm[0][0];
    const double p0 = 2.0 * sbzc::PI * sbzc::PI / (k * k * k) * P[SBZ_P_AS] * pow(k / P[SBZ_P_KPIVOT], P[SBZ_P_NS] - 1.0);
    pk[i] = p0 * D0 * D0;
    dlogpk[i] = 2.0 * (Dp - Dm) * tsc / D0;
}

// P(k, τ) = P0(k) Δ_S(τ)² for species combinations S (bit masks: c, b, h) at τ0 from the final
// states (taus null, n_t = 1) or at every saveat time from the saved states; a single species is
// its own Δ, several are ρ-weighted (spectrum_matter(modes, sol, k, τ) with
// total_symbolic_gauge_invariant_overdensities, fourier.jl:39-52,90-97).  out [n_m][n_t][n_inst].
extern "C" __global__ void sbz_pk_modes_kernel(const double* __restrict__ params, const double* __restrict__ knots,
                                               const long long* __restrict__ knot_off, const int* __restrict__ knot_n,
                                               const int* __restrict__ inst_set, const double* __restrict__ inst_k,
                                               const double* __restrict__ states, int n_inst, int n_t,
                                               const double* __restrict__ taus, int taus_stride,
                                               const int* __restrict__ masks, int n_m, double* __restrict__ out) {
    const long long e = blockIdx.x * (long long)blockDim.x + threadIdx.x;
    if (e >= (long long)n_inst * n_t) return;
    const int i = (int)(e / n_t), it = (int)(e % n_t);
    const int set = inst_set[i];
    const double* P = params + (size_t)set * SBZ_NPAR;
    const double* kt = knots + knot_off[set];
    const int nk = knot_n[set];
    const double k = inst_k[i];
    const double tau = taus ? taus[(size_t)set * taus_stride + it] : kt[nk - 1];
    double Cs[SBZ_NC], B[SBZ_NB];
    sbz_consts(P, Cs);
    sbz_consts_k(Cs, k);
    int idx = sbz_find_knot(kt, nk, tau);
    sbz_bg_plain(P, kt, nk, idx, tau, B);
    CSV u{states + ((size_t)i * n_t + it) * SBZ_N, 1};
    const double p0 = 2.0 * sbzc::PI * sbzc::PI / (k * k * k) * P[SBZ_P_AS] * pow(k / P[SBZ_P_KPIVOT], P[SBZ_P_NS] - 1.0);
    for (int m = 0; m < n_m; ++m) {
        const double d = sbz_dspecies(B, Cs, k, u, masks[m]);
        out[((size_t)m * n_t + it) * n_inst + i] = p0 * d * d;
    }
}

#endif  // SBZ_MULTI_FAMILY

#ifdef SBZ_HOST_PROBE
// Host entry points into the emitted model functions, compiled only into the test probe library
// (SBZ_DEV = __host__ __device__): they let the CPU test-suite compare the emitted RHS, W, LU,
// solves, ∂f/∂τ, initial conditions and Δ_m with the oracle without a GPU.
struct HV {
    double* p;
    double& operator[](int i) const { return p[i]; }
};
extern "C" int sbz_probe_n(void) { return SBZ_N; }
extern "C" int sbz_probe_nb(void) { return SBZ_NB; }
extern "C" int sbz_probe_nnz(void) { return SBZ_NNZ_LU; }
struct HP {  // parameter vector with its derived entries filled
    double p[SBZ_NPAR];
    explicit HP(const double* P) { for (int i = 0; i < SBZ_NPAR; ++i) p[i] = P[i]; sbz_params_derive(p); }
};
extern "C" void sbz_probe_bg(const double* P0, const double* kt, int nk, double tau, double* B, double* dB) {
    HP hp(P0);
    const double* P = hp.p;
    int idx = sbz_find_knot(kt, nk, tau);
    Dual Bd[SBZ_NB];
    sbz_pt_background<Dual>(P, kt, nk, idx, Dual(tau, 1.0), Bd);
    for (int i = 0; i < SBZ_NB; ++i) { B[i] = Bd[i].v; dB[i] = Bd[i].d; }
}
extern "C" void sbz_probe_src(const double* P0, const double* kt, int nk, double k, double tau, const double* u,
                              double* S, double* work) {
    HP hp(P0);
    const double* P = hp.p;
    double Cs[SBZ_NC]; sbz_consts(P, Cs);
    int idx = sbz_find_knot(kt, nk, tau);
    for (int i = 0; i < SBZ_N; ++i) work[i] = u[i];
    sbz_sources_at(P, kt, nk, idx, Cs, k, tau, HV{work}, HV{work + SBZ_N}, HV{work + 2 * SBZ_N},
                   HV{work + 3 * SBZ_N}, S);
}
extern "C" void sbz_probe_f(const double* B, const double* P, double k, const double* u, double* du) {
    double Cs[SBZ_NC]; sbz_consts(P, Cs);
    sbz_f(B, Cs, k, HV{(double*)u}, HV{du});
}
extern "C" void sbz_probe_w(const double* B, const double* P, double k, double ig, double* W) {
    double Cs[SBZ_NC]; sbz_consts(P, Cs);
    sbz_w(B, Cs, k, ig, HV{W});
}
extern "C" int sbz_probe_lu(double* W) { return sbz_lu(HV{W}) ? 1 : 0; }
extern "C" void sbz_probe_solve(const double* W, double* b) { sbz_solve(HV{(double*)W}, HV{b}); }
extern "C" void sbz_probe_dfdt(const double* B, const double* dB, const double* P, double k, const double* u, double* out) {
    double Cs[SBZ_NC]; sbz_consts(P, Cs);
    sbz_dfdt(B, dB, Cs, k, HV{(double*)u}, HV{out});
}
extern "C" void sbz_probe_ic(const double* B, const double* P, double k, double* u) {
    double Cs[SBZ_NC]; sbz_consts(P, Cs);
    sbz_ic(B, Cs, k, HV{u});
}
extern "C" double sbz_probe_dm(const double* B, const double* P, double k, const double* u) {
    double Cs[SBZ_NC]; sbz_consts(P, Cs);
    return sbz_dm(B, Cs, k, HV{(double*)u});
}
#endif
