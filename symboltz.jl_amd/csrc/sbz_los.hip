// Line-of-sight integration on the GPU (los_integrate, src/observables/angular.jl:109-185):
//   Θ_l(k) = Σ_τ w_τ S(τ, k) j_l(k (τ0 - τ)),  trapezoid weights w_τ over the τ grid,
// with j_l read from a uniform-x table by cubic Hermite interpolation in x (the reference's
// SphericalBesselCache, angular.jl:9-48).
//
// Layout: the Bessel table is [x][l] (l contiguous), a block handles one k and its threads walk
// l, so every table read of a wavefront is 64 consecutive doubles; the source column S(:, k) and
// the weights are staged through LDS in chunks and broadcast to all threads.  The kernel streams
// the table once per (k, τ) row: it is bound by HBM/L2 bandwidth, not arithmetic.
#include <hip/hip_runtime.h>

#include <cmath>

#include "sbz_internal.h"

#define LOS_CHUNK 256

extern "C" __global__ void __launch_bounds__(256)
sbz_los_kernel(const double* __restrict__ S, long long s_k, long long s_tau, const double* __restrict__ w,
               const double* __restrict__ chi, int n_tau, const double* __restrict__ ks,
               const double* __restrict__ y, const double* __restrict__ dy, int n_x, double dx, int n_l,
               double* __restrict__ out) {
    __shared__ double s_sw[LOS_CHUNK];
    __shared__ double s_chi[LOS_CHUNK];
    const int ik = blockIdx.y;
    const double k = ks[ik];
    const double invdx = 1.0 / dx;
    const int l0 = blockIdx.x * blockDim.x;
    const int il = l0 + threadIdx.x;
    double acc = 0.0;
    for (int c0 = 0; c0 < n_tau; c0 += LOS_CHUNK) {
        const int nc = min(LOS_CHUNK, n_tau - c0);
        __syncthreads();
        for (int j = threadIdx.x; j < nc; j += blockDim.x) {
            const int it = c0 + j;
            s_sw[j] = w[it] * S[(size_t)ik * s_k + (size_t)it * s_tau];
            s_chi[j] = chi[it];
        }
        __syncthreads();
        if (il < n_l) {
            for (int j = 0; j < nc; ++j) {
                const double x = k * s_chi[j];
                double fw = x * invdx;
                int i = (int)fw;
                if (i > n_x - 2) i = n_x - 2;  // the table spans [0, k_max χ_max]; clamp the last point
                fw -= (double)i;
                const double wm1 = fw - 1.0;
                const double* yr = y + (size_t)i * n_l + il;
                const double* dr = dy + (size_t)i * n_l + il;
                const double jl = (1.0 + 2.0 * fw) * wm1 * wm1 * yr[0] + fw * fw * (3.0 - 2.0 * fw) * yr[n_l] +
                                  fw * wm1 * (wm1 * dr[0] + fw * dr[n_l]) * dx;
                acc += s_sw[j] * jl;
            }
        }
    }
    if (il < n_l) out[(size_t)ik * n_l + il] = acc;
}

#define HIPCHK(x)                                                                                  \
    do {                                                                                           \
        hipError_t e_ = (x);                                                                       \
        if (e_ != hipSuccess) {                                                                    \
            sbz_set_error("%s failed: %s", #x, hipGetErrorString(e_));                            \
            st = SBZ_ERR_HIP;                                                                      \
            goto done;                                                                             \
        }                                                                                          \
    } while (0)

extern "C" int sbz_los_integrate(const double* S, long long s_k, long long s_tau, const double* taus, int n_tau,
                                 const double* ks, int n_k, const double* y, const double* dy, int n_x, double dx,
                                 int n_l, double* out, float* kernel_ms) {
    sbz_clear_error();
    if (!S || !taus || !ks || !y || !dy || !out || n_tau < 2 || n_k < 1 || n_l < 1 || n_x < 2 || !(dx > 0)) {
        sbz_set_error("sbz_los_integrate: invalid argument");
        return SBZ_ERR_ARG;
    }
    for (int i = 0; i < n_tau; ++i)
        if (!std::isfinite(taus[i]) || (i > 0 && !(taus[i] > taus[i - 1]))) {
            sbz_set_error("sbz_los_integrate: τs must be finite and sorted in ascending order");
            return SBZ_ERR_ARG;
        }
    double kmax = 0.0;
    for (int i = 0; i < n_k; ++i) {
        if (!(ks[i] > 0.0) || !(ks[i] < 1e300)) {
            sbz_set_error("sbz_los_integrate: wavenumber %d is not positive and finite", i);
            return SBZ_ERR_ARG;
        }
        kmax = ks[i] > kmax ? ks[i] : kmax;
    }
    if ((n_x - 1) * dx < kmax * (taus[n_tau - 1] - taus[0]) * (1.0 - 1e-12)) {
        sbz_set_error("sbz_los_integrate: Bessel table ends at x = %g < k_max χ_max = %g", (n_x - 1) * dx,
                      kmax * (taus[n_tau - 1] - taus[0]));
        return SBZ_ERR_ARG;
    }
    int st = SBZ_OK;
    const double tau0 = taus[n_tau - 1];
    double *w = nullptr, *chi = nullptr, *dS = nullptr, *dk = nullptr, *dyv = nullptr, *ddy = nullptr, *dout = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    size_t nS = (size_t)(n_k - 1) * s_k + (size_t)(n_tau - 1) * s_tau + 1;
    {
        // trapezoid weights (angular.jl:124-129) and χ = τ0 - τ
        double* hw = new double[n_tau];
        double* hc = new double[n_tau];
        hw[0] = 0.5 * (taus[1] - taus[0]);
        for (int i = 1; i < n_tau - 1; ++i) hw[i] = 0.5 * (taus[i + 1] - taus[i - 1]);
        hw[n_tau - 1] = 0.5 * (taus[n_tau - 1] - taus[n_tau - 2]);
        for (int i = 0; i < n_tau; ++i) hc[i] = tau0 - taus[i];
        hipError_t e = hipMalloc(&w, sizeof(double) * n_tau);
        if (e == hipSuccess) e = hipMalloc(&chi, sizeof(double) * n_tau);
        if (e == hipSuccess) e = hipMemcpy(w, hw, sizeof(double) * n_tau, hipMemcpyHostToDevice);
        if (e == hipSuccess) e = hipMemcpy(chi, hc, sizeof(double) * n_tau, hipMemcpyHostToDevice);
        delete[] hw;
        delete[] hc;
        HIPCHK(e);
    }
    HIPCHK(hipMalloc(&dS, sizeof(double) * nS));
    HIPCHK(hipMalloc(&dk, sizeof(double) * n_k));
    HIPCHK(hipMalloc(&dyv, sizeof(double) * (size_t)n_x * n_l));
    HIPCHK(hipMalloc(&ddy, sizeof(double) * (size_t)n_x * n_l));
    HIPCHK(hipMalloc(&dout, sizeof(double) * (size_t)n_k * n_l));
    HIPCHK(hipMemcpy(dS, S, sizeof(double) * nS, hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(dk, ks, sizeof(double) * n_k, hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(dyv, y, sizeof(double) * (size_t)n_x * n_l, hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(ddy, dy, sizeof(double) * (size_t)n_x * n_l, hipMemcpyHostToDevice));
    HIPCHK(hipEventCreate(&e0));
    HIPCHK(hipEventCreate(&e1));
    {
        dim3 block(256), grid((n_l + 255) / 256, n_k);
        HIPCHK(hipEventRecord(e0, 0));
        hipLaunchKernelGGL(sbz_los_kernel, grid, block, 0, 0, dS, s_k, s_tau, w, chi, n_tau, dk, dyv, ddy, n_x, dx,
                           n_l, dout);
        HIPCHK(hipGetLastError());
        HIPCHK(hipEventRecord(e1, 0));
        HIPCHK(hipEventSynchronize(e1));
        float ms = 0.0f;
        HIPCHK(hipEventElapsedTime(&ms, e0, e1));
        if (kernel_ms) *kernel_ms = ms;
    }
    HIPCHK(hipMemcpy(out, dout, sizeof(double) * (size_t)n_k * n_l, hipMemcpyDeviceToHost));
done:
    for (double* p : {w, chi, dS, dk, dyv, ddy, dout})
        if (p) (void)hipFree(p);
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    return st;
}
