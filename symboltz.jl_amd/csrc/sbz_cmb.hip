// CMB angular power spectra on the GPU: the stage of spectrum_cmb (src/observables/angular.jl:260-341)
// that follows the source solve, kept on the device from the coarse sources to C_l.
//
//   sbz_bessel_kernel   the SphericalBesselCache table (angular.jl:9-48): j_l(x) and
//                       j_l'(x) = l/(2l+1) j_{l-1} - (l+1)/(2l+1) j_{l+1} (:60) on x_i = i dx, one thread
//                       per x, by the three-term recurrence: upward from j_0, j_1 where every l is
//                       below x, else Miller's downward recurrence normalised to j_0 and j_1 (run
//                       twice: the first pass finds the normalisation and the rescaling exponent)
//   sbz_kinterp_kernel  source_kinterp with a ChebyshevInterpolator (fourier.jl:232-242,520-532):
//                       barycentric interpolation in y = f(k); a block takes 8 fine k, so every
//                       coarse source value it reads serves 8 outputs
//   sbz_cmb_los_kernel  los_integrate (angular.jl:109-185) of all source kinds at once: each
//                       Hermite read of the Bessel table serves every source
//   sbz_cmb_post_kernel Limber for l ≥ l_limber (angular.jl:150-177) and the T/E factors (:300-307)
//   sbz_cl_kernel       C_l = Σ_k w_k Θ_A Θ_B with w_k = q_k (2/π) k² P0(k), q the weights of the
//                       natural-cubic-spline integral over the fine k grid (angular.jl:198-223),
//                       as partial sums over k chunks reduced in a fixed order (deterministic)
//
// Layout: Bessel table [x][l] (l contiguous: the LOS threads walk l, so a wavefront reads 64
// consecutive doubles); sources [k][τ][s]; Θ [s][k][l].  The LOS kernel streams the table once per
// (k, τ): it is bound by L2/HBM bandwidth.
#include <hip/hip_runtime.h>

#include <climits>
#include <cmath>
#include <vector>

#include "sbz_internal.h"

#define SBZ_CMB_MAXS 4     // source kinds per call
#define SBZ_CMB_MAXC 512   // Chebyshev nodes per call
#define SBZ_KI_KB 8        // fine k per interpolation block
#define SBZ_LOS_CHUNK 256  // τ values staged per LDS pass
#define SBZ_CL_KCHUNK 512  // k per partial C_l sum

extern "C" __global__ void __launch_bounds__(256)
sbz_bessel_kernel(const int* __restrict__ ls, int n_l, int lmax, int n_x, double dx, double* __restrict__ y,
                  double* __restrict__ dy) {
    const long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x;
    if (i >= n_x) return;
    const long long ii = i < n_x - 1 ? i : n_x - 2;   // the last row duplicates the one before it
    const double x = (double)ii * dx;
    double* yr = y + i * n_l;
    double* dr = dy + i * n_l;
    if (x == 0.0) {
        for (int il = 0; il < n_l; ++il) {
            yr[il] = ls[il] == 0 ? 1.0 : 0.0;
            dr[il] = ls[il] == 1 ? 1.0 / 3.0 : 0.0;
        }
        return;
    }
    const double s = sin(x), c = cos(x);
    const double j0 = s / x, j1 = (s / x - c) / x;
    const double invx = 1.0 / x;
    const int L = lmax + 1;
    if (x > L + 1.0) {
        // upward: (jm, jc) = (j_{l-1}, j_l)
        int il = 0;
        while (il < n_l && ls[il] == 0) {
            yr[il] = j0;
            dr[il] = -j1;
            ++il;
        }
        double jm = j0, jc = j1;
        int next = il < n_l ? ls[il] : INT_MAX;
        for (int l = 1; l <= lmax && il < n_l; ++l) {
            const double jn = (2.0 * l + 1.0) * invx * jc - jm;
            while (l == next) {
                yr[il] = jc;
                dr[il] = (l * jm - (l + 1.0) * jn) / (2.0 * l + 1.0);
                ++il;
                next = il < n_l ? ls[il] : INT_MAX;
            }
            jm = jc;
            jc = jn;
        }
        return;
    }
    // Miller: start past the turning point max(L, x), where the ratio of the recurrence's two
    // solutions has decayed below rounding
    const double top = fmax((double)L, x);
    const int ltop = (int)(top + 40.0 + 12.0 * cbrt(top)) + 1;
    double C = 0.0;
    int ef = 0;
    for (int pass = 0; pass < 2; ++pass) {
        double dp = 0.0, dc = 1.0;   // (d_{l+1}, d_l)
        int e = 0;
        int il = n_l - 1;
        int next = il >= 0 ? ls[il] : -1;
        for (int l = ltop; l >= 1; --l) {
            const double dm = (2.0 * l + 1.0) * invx * dc - dp;
            if (pass == 1) {
                while (l == next) {
                    yr[il] = C * ldexp(dc, e - ef);
                    dr[il] = C * ldexp((l * dm - (l + 1.0) * dp) / (2.0 * l + 1.0), e - ef);
                    --il;
                    next = il >= 0 ? ls[il] : -1;
                }
            }
            dp = dc;
            dc = dm;
            if (fabs(dc) > 1e200) {
                dc = ldexp(dc, -600);
                dp = ldexp(dp, -600);
                e += 600;
            }
        }
        if (pass == 0) {
            const double m = fmax(fabs(dc), fabs(dp));
            const double a0 = dc / m, a1 = dp / m;
            C = (j0 * a0 + j1 * a1) / (a0 * a0 + a1 * a1) / m;
            ef = e;
        } else {
            while (il >= 0 && ls[il] == 0) {
                yr[il] = C * dc;
                dr[il] = -C * dp;
                --il;
            }
        }
    }
}

// S [n_c][m] on the nodes y_c (barycentric weights w_c) → out [n_f][m] at y_f
extern "C" __global__ void __launch_bounds__(256)
sbz_kinterp_kernel(const double* __restrict__ S, int n_c, int m, const double* __restrict__ yc,
                   const double* __restrict__ wc, const double* __restrict__ yf, int n_f, double* __restrict__ out) {
    __shared__ double tw[SBZ_KI_KB][SBZ_CMB_MAXC];
    __shared__ double inv_den[SBZ_KI_KB];
    __shared__ int hit[SBZ_KI_KB];
    const int f0 = blockIdx.x * SBZ_KI_KB;
    if (threadIdx.x < SBZ_KI_KB) {
        const int kk = threadIdx.x, f = f0 + kk;
        double den = 0.0;
        int h = -1;
        if (f < n_f) {
            const double y = yf[f];
            for (int j = 0; j < n_c; ++j) {
                const double d = y - yc[j];
                double t = 0.0;
                if (d == 0.0) {
                    if (h < 0) h = j;
                } else {
                    t = wc[j] / d;
                }
                tw[kk][j] = t;
                den += t;
            }
        }
        inv_den[kk] = den != 0.0 ? 1.0 / den : 0.0;
        hit[kk] = h;
    }
    __syncthreads();
    for (int q = threadIdx.x; q < m; q += blockDim.x) {
        double acc[SBZ_KI_KB];
#pragma unroll
        for (int kk = 0; kk < SBZ_KI_KB; ++kk) acc[kk] = 0.0;
        for (int j = 0; j < n_c; ++j) {
            const double s = S[(long long)j * m + q];
#pragma unroll
            for (int kk = 0; kk < SBZ_KI_KB; ++kk) acc[kk] += tw[kk][j] * s;
        }
#pragma unroll
        for (int kk = 0; kk < SBZ_KI_KB; ++kk) {
            const int f = f0 + kk;
            if (f < n_f)
                out[(long long)f * m + q] = hit[kk] >= 0 ? S[(long long)hit[kk] * m + q] : acc[kk] * inv_den[kk];
        }
    }
}

// Θ[s][k][l] = Σ_τ w_τ S[k][τ][s] j_l(k χ_τ) for l below that source's Limber index
extern "C" __global__ void __launch_bounds__(256)
sbz_cmb_los_kernel(const double* __restrict__ Sf, int n_s, const double* __restrict__ w, const double* __restrict__ chi,
                   int n_t, const double* __restrict__ ks, int n_f, const double* __restrict__ y,
                   const double* __restrict__ dy, int n_x, double dx, int n_l, int il_lim0, int il_lim1, int il_lim2,
                   int il_lim3, double* __restrict__ Th) {
    __shared__ double s_sw[SBZ_CMB_MAXS][SBZ_LOS_CHUNK];
    __shared__ double s_chi[SBZ_LOS_CHUNK];
    const int ik = blockIdx.y;
    const double k = ks[ik];
    const double invdx = 1.0 / dx;
    const int il = blockIdx.x * blockDim.x + threadIdx.x;
    const int lim[SBZ_CMB_MAXS] = {il_lim0, il_lim1, il_lim2, il_lim3};
    double acc[SBZ_CMB_MAXS] = {0.0, 0.0, 0.0, 0.0};
    for (int c0 = 0; c0 < n_t; c0 += SBZ_LOS_CHUNK) {
        const int nc = min(SBZ_LOS_CHUNK, n_t - c0);
        __syncthreads();
        for (int j = threadIdx.x; j < nc; j += blockDim.x) {
            const int it = c0 + j;
            for (int s = 0; s < n_s; ++s) s_sw[s][j] = w[it] * Sf[((long long)ik * n_t + it) * n_s + s];
            s_chi[j] = chi[it];
        }
        __syncthreads();
        if (il < n_l) {
            for (int j = 0; j < nc; ++j) {
                double fw = k * s_chi[j] * invdx;
                int i = (int)fw;
                if (i > n_x - 2) i = n_x - 2;
                fw -= (double)i;
                const double wm1 = fw - 1.0;
                const double* yr = y + (long long)i * n_l + il;
                const double* dr = dy + (long long)i * n_l + il;
                const double jl = (1.0 + 2.0 * fw) * wm1 * wm1 * yr[0] + fw * fw * (3.0 - 2.0 * fw) * yr[n_l] +
                                  fw * wm1 * (wm1 * dr[0] + fw * dr[n_l]) * dx;
#pragma unroll
                for (int s = 0; s < SBZ_CMB_MAXS; ++s)
                    if (s < n_s && il < lim[s]) acc[s] += s_sw[s][j] * jl;
            }
        }
    }
    if (il < n_l)
        for (int s = 0; s < n_s; ++s) Th[((long long)s * n_f + ik) * n_l + il] = acc[s];
}

// Limber entries and the per-kind factors, one thread per (k, l) and all sources
extern "C" __global__ void __launch_bounds__(256)
sbz_cmb_post_kernel(const double* __restrict__ Sf, int n_s, const double* __restrict__ taus,
                    const double* __restrict__ chi, int n_t, const double* __restrict__ ks, int n_f,
                    const int* __restrict__ ls, int n_l, const int* __restrict__ il_lim, const int* __restrict__ kind,
                    double* __restrict__ Th) {
    const long long e = blockIdx.x * (long long)blockDim.x + threadIdx.x;
    if (e >= (long long)n_f * n_l) return;
    const int ik = (int)(e / n_l), il = (int)(e % n_l);
    const double k = ks[ik];
    const int l = ls[il];
    const double tau0 = taus[n_t - 1];
    for (int s = 0; s < n_s; ++s) {
        double v = Th[((long long)s * n_f + ik) * n_l + il];
        if (il >= il_lim[s]) {
            v = 0.0;
            const double c = (l + 0.5) / k;
            if (c <= chi[0]) {
                // first index with τ ≥ τ0 - χ
                int lo = 0, hi = n_t;
                const double tt = tau0 - c;
                while (lo < hi) {
                    const int mid = (lo + hi) >> 1;
                    if (taus[mid] < tt) lo = mid + 1;
                    else hi = mid;
                }
                const int im = lo;
                if (im >= 1 && im < n_t) {
                    auto Sv = [&](int it) { return Sf[((long long)ik * n_t + it) * n_s + s]; };
                    const int ip = im - 1;
                    const double cm = chi[im], cp = chi[ip], Sm = Sv(im), Sp = Sv(ip);
                    const double Dc = cp - cm;
                    const double dSm = im <= n_t - 2 ? (Sv(im + 1) - Sp) / (chi[im + 1] - cp) : (Sp - Sm) / Dc;
                    const double dSp = ip >= 1 ? (Sm - Sv(im - 2)) / (cm - chi[im - 2]) : (Sp - Sm) / Dc;
                    const double t = (c - cm) / Dc, t2 = t * t, t3 = t2 * t;
                    const double Sx = (2 * t3 - 3 * t2 + 1) * Sm + (t3 - 2 * t2 + t) * Dc * dSm +
                                      (-2 * t3 + 3 * t2) * Sp + (t3 - t2) * Dc * dSp;
                    v = sqrt(M_PI / (2.0 * l + 1.0)) * Sx / k;
                }
            }
        }
        if (kind[s] == 0) v /= k;
        else if (kind[s] == 1) v *= sqrt((l + 2.0) * (l + 1.0) * (double)l * (l - 1.0)) / (k * k);
        Th[((long long)s * n_f + ik) * n_l + il] = v;
    }
}

// partial C_l over one k chunk: part[chunk][mode][l]
extern "C" __global__ void __launch_bounds__(256)
sbz_cl_kernel(const double* __restrict__ Th, int n_f, int n_l, const double* __restrict__ wk,
              const int* __restrict__ modes, int n_m, double* __restrict__ part) {
    __shared__ double red[4][64];
    const int c = threadIdx.x & 63, sl = threadIdx.x >> 6;
    const int il = blockIdx.x * 64 + c;
    const int k0 = blockIdx.y * SBZ_CL_KCHUNK, k1 = min(n_f, k0 + SBZ_CL_KCHUNK);
    for (int mo = 0; mo < n_m; ++mo) {
        const double* A = Th + (long long)modes[2 * mo] * n_f * n_l;
        const double* B = Th + (long long)modes[2 * mo + 1] * n_f * n_l;
        double acc = 0.0;
        if (il < n_l)
            for (int ik = k0 + sl; ik < k1; ik += 4)
                acc += wk[ik] * A[(long long)ik * n_l + il] * B[(long long)ik * n_l + il];
        red[sl][c] = acc;
        __syncthreads();
        if (sl == 0 && il < n_l)
            part[((long long)blockIdx.y * n_m + mo) * n_l + il] = (red[0][c] + red[1][c]) + (red[2][c] + red[3][c]);
        __syncthreads();
    }
}

extern "C" __global__ void sbz_cl_reduce_kernel(const double* __restrict__ part, int n_chunk, int n_ml,
                                                double* __restrict__ cl) {
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= n_ml) return;
    double s = 0.0;
    for (int c = 0; c < n_chunk; ++c) s += part[(long long)c * n_ml + e];
    cl[e] = s;
}

// ------------------------------------------------------------------------------------ host side

namespace {

struct DevBuf {
    std::vector<void*> ptrs;
    ~DevBuf() {
        for (void* p : ptrs) (void)hipFree(p);
    }
    template <class T>
    hipError_t alloc(T** p, size_t n) {
        hipError_t e = hipMalloc((void**)p, sizeof(T) * (n ? n : 1));
        if (e == hipSuccess) ptrs.push_back((void*)*p);
        return e;
    }
};

struct Events {
    std::vector<hipEvent_t> ev;
    ~Events() {
        for (auto e : ev) (void)hipEventDestroy(e);
    }
};

int check_ls(const int* ls, int n_l, const char* who) {
    for (int i = 0; i < n_l; ++i)
        if (ls[i] < 0 || (i > 0 && ls[i] <= ls[i - 1])) {
            sbz_set_error("%s: ls must be non-negative and strictly ascending (entry %d)", who, i);
            return SBZ_ERR_ARG;
        }
    return SBZ_OK;
}

}  // namespace

#define CCHK(x)                                                                                   \
    do {                                                                                          \
        hipError_t e_ = (x);                                                                      \
        if (e_ != hipSuccess) {                                                                   \
            sbz_set_error("%s failed: %s", #x, hipGetErrorString(e_));                           \
            return SBZ_ERR_HIP;                                                                   \
        }                                                                                         \
    } while (0)

static int bessel_launch(const int* d_ls, int n_l, int lmax, int n_x, double dx, double* d_y, double* d_dy,
                         hipStream_t st) {
    hipLaunchKernelGGL(sbz_bessel_kernel, dim3((n_x + 255) / 256), dim3(256), 0, st, d_ls, n_l, lmax, n_x, dx, d_y,
                       d_dy);
    CCHK(hipGetLastError());
    return SBZ_OK;
}

extern "C" int sbz_bessel_table(const int* ls, int n_l, int n_x, double dx, double* y, double* dy) {
    sbz_clear_error();
    if (!ls || !y || !dy || n_l < 1 || n_x < 2 || !(dx > 0)) {
        sbz_set_error("sbz_bessel_table: invalid argument");
        return SBZ_ERR_ARG;
    }
    if (int st = check_ls(ls, n_l, "sbz_bessel_table")) return st;
    DevBuf b;
    int* d_ls;
    double *d_y, *d_dy;
    CCHK(b.alloc(&d_ls, n_l));
    CCHK(b.alloc(&d_y, (size_t)n_x * n_l));
    CCHK(b.alloc(&d_dy, (size_t)n_x * n_l));
    CCHK(hipMemcpy(d_ls, ls, sizeof(int) * n_l, hipMemcpyHostToDevice));
    if (int st = bessel_launch(d_ls, n_l, ls[n_l - 1], n_x, dx, d_y, d_dy, 0)) return st;
    CCHK(hipMemcpy(y, d_y, sizeof(double) * n_x * n_l, hipMemcpyDeviceToHost));
    CCHK(hipMemcpy(dy, d_dy, sizeof(double) * n_x * n_l, hipMemcpyDeviceToHost));
    return SBZ_OK;
}

extern "C" int sbz_cmb_spectra(const double* S_coarse, int n_c, int n_s, const double* taus, int n_t, const double* y_c,
                               const double* w_c, const double* ks, const double* y_f, const double* wk, int n_f,
                               const int* ls, int n_l, int n_x, double dx, const int* kind, const int* l_limber,
                               const int* modes, int n_m, double* out_cl, double* out_theta, float* kernel_ms) {
    sbz_clear_error();
    if (!S_coarse || !taus || !y_c || !w_c || !ks || !y_f || !wk || !ls || !kind || !l_limber || !modes || !out_cl ||
        n_c < 1 || n_c > SBZ_CMB_MAXC || n_s < 1 || n_s > SBZ_CMB_MAXS || n_t < 2 || n_f < 1 || n_l < 1 || n_x < 2 ||
        !(dx > 0) || n_m < 1) {
        sbz_set_error("sbz_cmb_spectra: invalid argument (at most %d nodes and %d source kinds)", SBZ_CMB_MAXC,
                      SBZ_CMB_MAXS);
        return SBZ_ERR_ARG;
    }
    if (int st = check_ls(ls, n_l, "sbz_cmb_spectra")) return st;
    if (n_f > 65535) {
        sbz_set_error("sbz_cmb_spectra: at most 65535 fine wavenumbers per call (got %d)", n_f);
        return SBZ_ERR_ARG;
    }
    for (int i = 0; i < n_t; ++i)
        if (!std::isfinite(taus[i]) || (i > 0 && !(taus[i] > taus[i - 1]))) {
            sbz_set_error("sbz_cmb_spectra: τs must be finite and sorted in ascending order");
            return SBZ_ERR_ARG;
        }
    for (int i = 0; i < n_c; ++i)
        if (!std::isfinite(y_c[i]) || !std::isfinite(w_c[i])) {
            sbz_set_error("sbz_cmb_spectra: non-finite interpolation node or weight %d", i);
            return SBZ_ERR_ARG;
        }
    for (int i = 0; i < n_f; ++i)
        if (!std::isfinite(y_f[i]) || !std::isfinite(wk[i])) {
            sbz_set_error("sbz_cmb_spectra: non-finite transformed wavenumber or weight %d", i);
            return SBZ_ERR_ARG;
        }
    double kmax = 0.0;
    for (int i = 0; i < n_f; ++i) {
        if (!(ks[i] > 0.0) || !(ks[i] < 1e300)) {
            sbz_set_error("sbz_cmb_spectra: wavenumber %d is not positive and finite", i);
            return SBZ_ERR_ARG;
        }
        kmax = fmax(kmax, ks[i]);
    }
    const double chimax = taus[n_t - 1] - taus[0];
    if ((n_x - 2) * dx < kmax * chimax * (1.0 - 1e-12)) {
        sbz_set_error("sbz_cmb_spectra: Bessel table ends at x = %g < k_max χ_max = %g", (n_x - 2) * dx,
                      kmax * chimax);
        return SBZ_ERR_ARG;
    }
    for (int s = 0; s < n_s; ++s)
        if (kind[s] < 0 || kind[s] > 2) {
            sbz_set_error("sbz_cmb_spectra: source kind %d of source %d is not 0 (T), 1 (E) or 2 (lensing)", kind[s],
                          s);
            return SBZ_ERR_ARG;
        }
    for (int mo = 0; mo < 2 * n_m; ++mo)
        if (modes[mo] < 0 || modes[mo] >= n_s) {
            sbz_set_error("sbz_cmb_spectra: mode entry %d names source %d of %d", mo, modes[mo], n_s);
            return SBZ_ERR_ARG;
        }
    const long long nS = (long long)n_c * n_t * n_s;
    for (long long e = 0; e < nS; ++e)
        if (!std::isfinite(S_coarse[e])) {
            sbz_set_error("sbz_cmb_spectra: non-finite source value at flat index %lld", e);
            return SBZ_ERR_ARG;
        }
    // first l index at or above each source's Limber multipole
    int il_lim[SBZ_CMB_MAXS] = {n_l, n_l, n_l, n_l};
    for (int s = 0; s < n_s; ++s) {
        int i = 0;
        while (i < n_l && ls[i] < l_limber[s]) ++i;
        il_lim[s] = i;
    }
    std::vector<double> w(n_t), chi(n_t);
    w[0] = 0.5 * (taus[1] - taus[0]);
    for (int i = 1; i < n_t - 1; ++i) w[i] = 0.5 * (taus[i + 1] - taus[i - 1]);
    w[n_t - 1] = 0.5 * (taus[n_t - 1] - taus[n_t - 2]);
    for (int i = 0; i < n_t; ++i) chi[i] = taus[n_t - 1] - taus[i];
    const int n_chunk = (n_f + SBZ_CL_KCHUNK - 1) / SBZ_CL_KCHUNK;

    DevBuf b;
    double *d_S, *d_yc, *d_wc, *d_yf, *d_ks, *d_wk, *d_taus, *d_w, *d_chi, *d_y, *d_dy, *d_Sf, *d_Th, *d_part, *d_cl;
    int *d_ls, *d_lim, *d_kind, *d_modes;
    const int m = n_t * n_s;
    CCHK(b.alloc(&d_S, (size_t)nS));
    CCHK(b.alloc(&d_yc, n_c));
    CCHK(b.alloc(&d_wc, n_c));
    CCHK(b.alloc(&d_yf, n_f));
    CCHK(b.alloc(&d_ks, n_f));
    CCHK(b.alloc(&d_wk, n_f));
    CCHK(b.alloc(&d_taus, n_t));
    CCHK(b.alloc(&d_w, n_t));
    CCHK(b.alloc(&d_chi, n_t));
    CCHK(b.alloc(&d_y, (size_t)n_x * n_l));
    CCHK(b.alloc(&d_dy, (size_t)n_x * n_l));
    CCHK(b.alloc(&d_Sf, (size_t)n_f * m));
    CCHK(b.alloc(&d_Th, (size_t)n_s * n_f * n_l));
    CCHK(b.alloc(&d_part, (size_t)n_chunk * n_m * n_l));
    CCHK(b.alloc(&d_cl, (size_t)n_m * n_l));
    CCHK(b.alloc(&d_ls, n_l));
    CCHK(b.alloc(&d_lim, n_s));
    CCHK(b.alloc(&d_kind, n_s));
    CCHK(b.alloc(&d_modes, 2 * n_m));
    CCHK(hipMemcpy(d_S, S_coarse, sizeof(double) * nS, hipMemcpyHostToDevice));
    CCHK(hipMemcpy(d_yc, y_c, sizeof(double) * n_c, hipMemcpyHostToDevice));
    CCHK(hipMemcpy(d_wc, w_c, sizeof(double) * n_c, hipMemcpyHostToDevice));
    CCHK(hipMemcpy(d_yf, y_f, sizeof(double) * n_f, hipMemcpyHostToDevice));
    CCHK(hipMemcpy(d_ks, ks, sizeof(double) * n_f, hipMemcpyHostToDevice));
    CCHK(hipMemcpy(d_wk, wk, sizeof(double) * n_f, hipMemcpyHostToDevice));
    CCHK(hipMemcpy(d_taus, taus, sizeof(double) * n_t, hipMemcpyHostToDevice));
    CCHK(hipMemcpy(d_w, w.data(), sizeof(double) * n_t, hipMemcpyHostToDevice));
    CCHK(hipMemcpy(d_chi, chi.data(), sizeof(double) * n_t, hipMemcpyHostToDevice));
    CCHK(hipMemcpy(d_ls, ls, sizeof(int) * n_l, hipMemcpyHostToDevice));
    CCHK(hipMemcpy(d_lim, il_lim, sizeof(int) * n_s, hipMemcpyHostToDevice));
    CCHK(hipMemcpy(d_kind, kind, sizeof(int) * n_s, hipMemcpyHostToDevice));
    CCHK(hipMemcpy(d_modes, modes, sizeof(int) * 2 * n_m, hipMemcpyHostToDevice));

    Events E;
    E.ev.resize(6, nullptr);
    for (auto& e : E.ev) CCHK(hipEventCreate(&e));
    CCHK(hipEventRecord(E.ev[0], 0));
    if (int st = bessel_launch(d_ls, n_l, ls[n_l - 1], n_x, dx, d_y, d_dy, 0)) return st;
    CCHK(hipEventRecord(E.ev[1], 0));
    hipLaunchKernelGGL(sbz_kinterp_kernel, dim3((n_f + SBZ_KI_KB - 1) / SBZ_KI_KB), dim3(256), 0, 0, d_S, n_c, m, d_yc,
                       d_wc, d_yf, n_f, d_Sf);
    CCHK(hipGetLastError());
    CCHK(hipEventRecord(E.ev[2], 0));
    {
        const int bl = n_l >= 256 ? 256 : ((n_l + 63) / 64) * 64;
        hipLaunchKernelGGL(sbz_cmb_los_kernel, dim3((n_l + bl - 1) / bl, n_f), dim3(bl), 0, 0, d_Sf, n_s, d_w, d_chi,
                           n_t, d_ks, n_f, d_y, d_dy, n_x, dx, n_l, il_lim[0], il_lim[1], il_lim[2], il_lim[3], d_Th);
        CCHK(hipGetLastError());
    }
    CCHK(hipEventRecord(E.ev[3], 0));
    {
        const long long n = (long long)n_f * n_l;
        hipLaunchKernelGGL(sbz_cmb_post_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, 0, d_Sf, n_s, d_taus,
                           d_chi, n_t, d_ks, n_f, d_ls, n_l, d_lim, d_kind, d_Th);
        CCHK(hipGetLastError());
    }
    CCHK(hipEventRecord(E.ev[4], 0));
    hipLaunchKernelGGL(sbz_cl_kernel, dim3((n_l + 63) / 64, n_chunk), dim3(256), 0, 0, d_Th, n_f, n_l, d_wk, d_modes,
                       n_m, d_part);
    CCHK(hipGetLastError());
    hipLaunchKernelGGL(sbz_cl_reduce_kernel, dim3((n_m * n_l + 255) / 256), dim3(256), 0, 0, d_part, n_chunk, n_m * n_l,
                       d_cl);
    CCHK(hipGetLastError());
    CCHK(hipEventRecord(E.ev[5], 0));
    CCHK(hipEventSynchronize(E.ev[5]));
    if (kernel_ms)
        for (int i = 0; i < 5; ++i) CCHK(hipEventElapsedTime(&kernel_ms[i], E.ev[i], E.ev[i + 1]));
    CCHK(hipMemcpy(out_cl, d_cl, sizeof(double) * n_m * n_l, hipMemcpyDeviceToHost));
    if (out_theta) CCHK(hipMemcpy(out_theta, d_Th, sizeof(double) * n_s * n_f * n_l, hipMemcpyDeviceToHost));
    return SBZ_OK;
}
