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
#include <cstring>
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

int check_ls(const int* ls, int n_l, const char* who) {
    for (int i = 0; i < n_l; ++i)
        if (ls[i] < 0 || (i > 0 && ls[i] <= ls[i - 1])) {
            sbz_set_error("%s: ls must be non-negative and strictly ascending (entry %d)", who, i);
            return SBZ_ERR_ARG;
        }
    return SBZ_OK;
}

// grow-only device buffer
struct DBuf {
    void* p = nullptr;
    size_t bytes = 0;
    hipError_t reserve(size_t b) {
        if (b <= bytes) return hipSuccess;
        if (p) (void)hipFree(p);
        p = nullptr;
        bytes = 0;
        hipError_t e = hipMalloc(&p, b);
        if (e == hipSuccess) bytes = b;
        return e;
    }
    template <class T>
    T* as() const {
        return (T*)p;
    }
    void release() {
        if (p) (void)hipFree(p);
        p = nullptr;
        bytes = 0;
    }
};

}  // namespace

#define CCHK(x)                                                                                   \
    do {                                                                                          \
        hipError_t e_ = (x);                                                                      \
        if (e_ != hipSuccess) {                                                                   \
            sbz_set_error("%s failed: %s", #x, hipGetErrorString(e_));                           \
            return SBZ_ERR_HIP;                                                                   \
        }                                                                                         \
    } while (0)

// A C_l context: the Bessel table of one l list stays on the device across calls, work buffers
// only grow, and every launch goes to the context's own stream.
struct sbz_cmb {
    int n_l = 0, n_x = 0, lmax = 0;
    double dx = 0.0;
    float bessel_ms = 0.0f;      // the table kernel's time at creation
    std::vector<int> ls;
    hipStream_t stream = nullptr;
    hipEvent_t ev[6] = {};
    DBuf ls_d, y, dy;            // the table
    DBuf S, Sf, Th, part, cl;    // per call
    DBuf small;                  // nodes, weights, grids, small int arrays
    ~sbz_cmb() {
        for (DBuf* b : {&ls_d, &y, &dy, &S, &Sf, &Th, &part, &cl, &small}) b->release();
        for (auto e : ev)
            if (e) (void)hipEventDestroy(e);
        if (stream) (void)hipStreamDestroy(stream);
    }
};

static int bessel_launch(const int* d_ls, int n_l, int lmax, int n_x, double dx, double* d_y, double* d_dy,
                         hipStream_t st) {
    hipLaunchKernelGGL(sbz_bessel_kernel, dim3((n_x + 255) / 256), dim3(256), 0, st, d_ls, n_l, lmax, n_x, dx, d_y,
                       d_dy);
    CCHK(hipGetLastError());
    return SBZ_OK;
}

extern "C" int sbz_cmb_create(const int* ls, int n_l, int n_x, double dx, sbz_cmb** out) {
    sbz_clear_error();
    if (!ls || !out || n_l < 1 || n_x < 3 || !(dx > 0) || !std::isfinite(dx)) {
        sbz_set_error("sbz_cmb_create: invalid argument");
        return SBZ_ERR_ARG;
    }
    if (int st = check_ls(ls, n_l, "sbz_cmb_create")) return st;
    sbz_cmb* h = new sbz_cmb();
    h->n_l = n_l;
    h->n_x = n_x;
    h->dx = dx;
    h->lmax = ls[n_l - 1];
    h->ls.assign(ls, ls + n_l);
    auto fail = [&](hipError_t e, const char* what) {
        sbz_set_error("sbz_cmb_create: %s: %s", what, hipGetErrorString(e));
        delete h;
        return SBZ_ERR_HIP;
    };
    hipError_t e = hipStreamCreate(&h->stream);
    if (e != hipSuccess) return fail(e, "stream");
    for (auto& v : h->ev)
        if ((e = hipEventCreate(&v)) != hipSuccess) return fail(e, "event");
    const size_t tb = sizeof(double) * (size_t)n_x * n_l;
    if ((e = h->ls_d.reserve(sizeof(int) * n_l)) != hipSuccess) return fail(e, "hipMalloc");
    if ((e = h->y.reserve(tb)) != hipSuccess || (e = h->dy.reserve(tb)) != hipSuccess) return fail(e, "hipMalloc");
    if ((e = hipMemcpyAsync(h->ls_d.p, ls, sizeof(int) * n_l, hipMemcpyHostToDevice, h->stream)) != hipSuccess)
        return fail(e, "upload");
    if ((e = hipEventRecord(h->ev[0], h->stream)) != hipSuccess) return fail(e, "event");
    hipLaunchKernelGGL(sbz_bessel_kernel, dim3((n_x + 255) / 256), dim3(256), 0, h->stream, h->ls_d.as<int>(), n_l,
                       h->lmax, n_x, dx, h->y.as<double>(), h->dy.as<double>());
    if ((e = hipGetLastError()) != hipSuccess) return fail(e, "launch");
    if ((e = hipEventRecord(h->ev[1], h->stream)) != hipSuccess) return fail(e, "event");
    if ((e = hipStreamSynchronize(h->stream)) != hipSuccess) return fail(e, "sbz_bessel_kernel");
    (void)hipEventElapsedTime(&h->bessel_ms, h->ev[0], h->ev[1]);
    *out = h;
    return SBZ_OK;
}

extern "C" int sbz_cmb_destroy(sbz_cmb* h) {
    delete h;
    return SBZ_OK;
}

extern "C" int sbz_bessel_table(const int* ls, int n_l, int n_x, double dx, double* y, double* dy) {
    sbz_clear_error();
    if (!y || !dy) {
        sbz_set_error("sbz_bessel_table: invalid argument");
        return SBZ_ERR_ARG;
    }
    sbz_cmb* h = nullptr;
    if (int st = sbz_cmb_create(ls, n_l, n_x, dx, &h)) return st;
    int st = SBZ_OK;
    hipError_t e = hipMemcpy(y, h->y.p, sizeof(double) * n_x * n_l, hipMemcpyDeviceToHost);
    if (e == hipSuccess) e = hipMemcpy(dy, h->dy.p, sizeof(double) * n_x * n_l, hipMemcpyDeviceToHost);
    if (e != hipSuccess) {
        sbz_set_error("sbz_bessel_table: copy: %s", hipGetErrorString(e));
        st = SBZ_ERR_HIP;
    }
    sbz_cmb_destroy(h);
    return st;
}

extern "C" int sbz_cmb_compute(sbz_cmb* h, const double* S_coarse, int n_c, int n_s, const double* taus, int n_t,
                               const double* y_c, const double* w_c, const double* ks, const double* y_f,
                               const double* wk, int n_f, const int* kind, const int* l_limber, const int* modes,
                               int n_m, double* out_cl, double* out_theta, float* kernel_ms) {
    sbz_clear_error();
    if (!h || !S_coarse || !taus || !y_c || !w_c || !ks || !y_f || !wk || !kind || !l_limber || !modes || !out_cl ||
        n_c < 1 || n_c > SBZ_CMB_MAXC || n_s < 1 || n_s > SBZ_CMB_MAXS || n_t < 2 || n_f < 1 || n_m < 1) {
        sbz_set_error("sbz_cmb_compute: invalid argument (at most %d nodes and %d source kinds)", SBZ_CMB_MAXC,
                      SBZ_CMB_MAXS);
        return SBZ_ERR_ARG;
    }
    const int n_l = h->n_l, n_x = h->n_x;
    const double dx = h->dx;
    if (n_f > 65535) {
        sbz_set_error("sbz_cmb_compute: at most 65535 fine wavenumbers per call (got %d)", n_f);
        return SBZ_ERR_ARG;
    }
    for (int i = 0; i < n_t; ++i)
        if (!std::isfinite(taus[i]) || (i > 0 && !(taus[i] > taus[i - 1]))) {
            sbz_set_error("sbz_cmb_compute: τs must be finite and sorted in ascending order");
            return SBZ_ERR_ARG;
        }
    for (int i = 0; i < n_c; ++i)
        if (!std::isfinite(y_c[i]) || !std::isfinite(w_c[i])) {
            sbz_set_error("sbz_cmb_compute: non-finite interpolation node or weight %d", i);
            return SBZ_ERR_ARG;
        }
    double kmax = 0.0;
    for (int i = 0; i < n_f; ++i) {
        if (!(ks[i] > 0.0) || !(ks[i] < 1e300) || !std::isfinite(y_f[i]) || !std::isfinite(wk[i])) {
            sbz_set_error("sbz_cmb_compute: wavenumber %d (or its transform or weight) is not positive and finite", i);
            return SBZ_ERR_ARG;
        }
        kmax = fmax(kmax, ks[i]);
    }
    const double chimax = taus[n_t - 1] - taus[0];
    if ((n_x - 2) * dx < kmax * chimax * (1.0 - 1e-12)) {
        sbz_set_error("sbz_cmb_compute: Bessel table ends at x = %g < k_max χ_max = %g", (n_x - 2) * dx,
                      kmax * chimax);
        return SBZ_ERR_ARG;
    }
    for (int s = 0; s < n_s; ++s)
        if (kind[s] < 0 || kind[s] > 2) {
            sbz_set_error("sbz_cmb_compute: source kind %d of source %d is not 0 (T), 1 (E) or 2 (lensing)", kind[s],
                          s);
            return SBZ_ERR_ARG;
        }
    for (int mo = 0; mo < 2 * n_m; ++mo)
        if (modes[mo] < 0 || modes[mo] >= n_s) {
            sbz_set_error("sbz_cmb_compute: mode entry %d names source %d of %d", mo, modes[mo], n_s);
            return SBZ_ERR_ARG;
        }
    const long long nS = (long long)n_c * n_t * n_s;
    for (long long e = 0; e < nS; ++e)
        if (!std::isfinite(S_coarse[e])) {
            sbz_set_error("sbz_cmb_compute: non-finite source value at flat index %lld", e);
            return SBZ_ERR_ARG;
        }
    // first l index at or above each source's Limber multipole
    int il_lim[SBZ_CMB_MAXS] = {n_l, n_l, n_l, n_l};
    for (int s = 0; s < n_s; ++s) {
        int i = 0;
        while (i < n_l && h->ls[i] < l_limber[s]) ++i;
        il_lim[s] = i;
    }
    const int n_chunk = (n_f + SBZ_CL_KCHUNK - 1) / SBZ_CL_KCHUNK;
    const int m = n_t * n_s;
    // small arrays packed into one upload: yc wc | yf ks wk | taus w chi | ints (lim kind modes)
    std::vector<double> sm;
    sm.reserve(2 * n_c + 3 * n_f + 3 * n_t + SBZ_CMB_MAXS * 2 + 2 * n_m + 8);
    const size_t o_yc = sm.size();
    sm.insert(sm.end(), y_c, y_c + n_c);
    const size_t o_wc = sm.size();
    sm.insert(sm.end(), w_c, w_c + n_c);
    const size_t o_yf = sm.size();
    sm.insert(sm.end(), y_f, y_f + n_f);
    const size_t o_ks = sm.size();
    sm.insert(sm.end(), ks, ks + n_f);
    const size_t o_wk = sm.size();
    sm.insert(sm.end(), wk, wk + n_f);
    const size_t o_taus = sm.size();
    sm.insert(sm.end(), taus, taus + n_t);
    // trapezoid weights (angular.jl:124-129) and χ = τ0 - τ
    const size_t o_w = sm.size();
    sm.resize(sm.size() + n_t);
    sm[o_w] = 0.5 * (taus[1] - taus[0]);
    for (int i = 1; i < n_t - 1; ++i) sm[o_w + i] = 0.5 * (taus[i + 1] - taus[i - 1]);
    sm[o_w + n_t - 1] = 0.5 * (taus[n_t - 1] - taus[n_t - 2]);
    const size_t o_chi = sm.size();
    for (int i = 0; i < n_t; ++i) sm.push_back(taus[n_t - 1] - taus[i]);
    const size_t o_int = sm.size();   // ints stored after the doubles
    std::vector<int> iv;
    for (int s = 0; s < n_s; ++s) iv.push_back(il_lim[s]);
    for (int s = 0; s < n_s; ++s) iv.push_back(kind[s]);
    for (int q = 0; q < 2 * n_m; ++q) iv.push_back(modes[q]);
    sm.resize(o_int + (iv.size() + 1) / 2);
    std::memcpy(sm.data() + o_int, iv.data(), sizeof(int) * iv.size());

    CCHK(h->small.reserve(sizeof(double) * sm.size()));
    CCHK(h->S.reserve(sizeof(double) * nS));
    CCHK(h->Sf.reserve(sizeof(double) * (size_t)n_f * m));
    CCHK(h->Th.reserve(sizeof(double) * (size_t)n_s * n_f * n_l));
    CCHK(h->part.reserve(sizeof(double) * (size_t)n_chunk * n_m * n_l));
    CCHK(h->cl.reserve(sizeof(double) * (size_t)n_m * n_l));
    hipStream_t st = h->stream;
    double* dsm = h->small.as<double>();
    const int* d_lim = (const int*)(dsm + o_int);
    const int* d_kind = d_lim + n_s;
    const int* d_modes = d_kind + n_s;
    CCHK(hipMemcpyAsync(dsm, sm.data(), sizeof(double) * sm.size(), hipMemcpyHostToDevice, st));
    CCHK(hipMemcpyAsync(h->S.p, S_coarse, sizeof(double) * nS, hipMemcpyHostToDevice, st));
    CCHK(hipEventRecord(h->ev[0], st));
    CCHK(hipEventRecord(h->ev[1], st));   // the table is built at creation
    hipLaunchKernelGGL(sbz_kinterp_kernel, dim3((n_f + SBZ_KI_KB - 1) / SBZ_KI_KB), dim3(256), 0, st,
                       h->S.as<double>(), n_c, m, dsm + o_yc, dsm + o_wc, dsm + o_yf, n_f, h->Sf.as<double>());
    CCHK(hipGetLastError());
    CCHK(hipEventRecord(h->ev[2], st));
    {
        const int bl = n_l >= 256 ? 256 : ((n_l + 63) / 64) * 64;
        hipLaunchKernelGGL(sbz_cmb_los_kernel, dim3((n_l + bl - 1) / bl, n_f), dim3(bl), 0, st, h->Sf.as<double>(),
                           n_s, dsm + o_w, dsm + o_chi, n_t, dsm + o_ks, n_f, h->y.as<double>(), h->dy.as<double>(),
                           n_x, dx, n_l, il_lim[0], il_lim[1], il_lim[2], il_lim[3], h->Th.as<double>());
        CCHK(hipGetLastError());
    }
    CCHK(hipEventRecord(h->ev[3], st));
    {
        const long long n = (long long)n_f * n_l;
        hipLaunchKernelGGL(sbz_cmb_post_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st,
                           h->Sf.as<double>(), n_s, dsm + o_taus, dsm + o_chi, n_t, dsm + o_ks, n_f, h->ls_d.as<int>(),
                           n_l, d_lim, d_kind, h->Th.as<double>());
        CCHK(hipGetLastError());
    }
    CCHK(hipEventRecord(h->ev[4], st));
    hipLaunchKernelGGL(sbz_cl_kernel, dim3((n_l + 63) / 64, n_chunk), dim3(256), 0, st, h->Th.as<double>(), n_f, n_l,
                       dsm + o_wk, d_modes, n_m, h->part.as<double>());
    CCHK(hipGetLastError());
    hipLaunchKernelGGL(sbz_cl_reduce_kernel, dim3((n_m * n_l + 255) / 256), dim3(256), 0, st, h->part.as<double>(),
                       n_chunk, n_m * n_l, h->cl.as<double>());
    CCHK(hipGetLastError());
    CCHK(hipEventRecord(h->ev[5], st));
    CCHK(hipMemcpyAsync(out_cl, h->cl.p, sizeof(double) * n_m * n_l, hipMemcpyDeviceToHost, st));
    if (out_theta)
        CCHK(hipMemcpyAsync(out_theta, h->Th.p, sizeof(double) * n_s * n_f * n_l, hipMemcpyDeviceToHost, st));
    CCHK(hipStreamSynchronize(st));
    if (kernel_ms)
        for (int i = 0; i < 5; ++i) CCHK(hipEventElapsedTime(&kernel_ms[i], h->ev[i], h->ev[i + 1]));
    return SBZ_OK;
}

extern "C" int sbz_cmb_spectra(const double* S_coarse, int n_c, int n_s, const double* taus, int n_t, const double* y_c,
                               const double* w_c, const double* ks, const double* y_f, const double* wk, int n_f,
                               const int* ls, int n_l, int n_x, double dx, const int* kind, const int* l_limber,
                               const int* modes, int n_m, double* out_cl, double* out_theta, float* kernel_ms) {
    sbz_cmb* h = nullptr;
    int st = sbz_cmb_create(ls, n_l, n_x, dx, &h);
    if (st == SBZ_OK)
        st = sbz_cmb_compute(h, S_coarse, n_c, n_s, taus, n_t, y_c, w_c, ks, y_f, wk, n_f, kind, l_limber, modes, n_m,
                             out_cl, out_theta, kernel_ms);
    if (st == SBZ_OK && kernel_ms) kernel_ms[0] = h->bessel_ms;
    if (h) sbz_cmb_destroy(h);
    return st;
}
