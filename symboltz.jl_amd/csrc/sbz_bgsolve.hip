// Batched background + thermodynamics solve on the GPU: one thread integrates one parameter set
// (the shared solver of sbz_bgsolve.h: Rodas5P on the five background unknowns with the a = 1
// event, solve.jl:158-203,427-435).  The reference solves the background of each cosmology
// serially on the CPU before its perturbations; for many parameter sets (Latin-hypercube batches,
// surrogate training data) the sets run here side by side, 64 per wavefront.
//
// Layout: params [n][SBZ_NPAR] (in/out), knots t [n][max_knots], u and du [n][max_knots][5].
#include <hip/hip_runtime.h>

#include <cmath>
#include <vector>

#include "sbz_bgsolve.h"
#include "sbz_internal.h"

extern "C" __global__ void __launch_bounds__(64)
sbz_bg_solve_kernel(double* __restrict__ params, int n, double rtol, double atol, int max_knots,
                    double* __restrict__ t_out, double* __restrict__ u_out, double* __restrict__ du_out,
                    int* __restrict__ n_knots, int* __restrict__ n_rejected, int* __restrict__ status,
                    double* __restrict__ t_fail) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const size_t o = (size_t)i * max_knots;
    status[i] = sbz_bg_solve(params + (size_t)i * SBZ_NPAR, rtol, atol, max_knots, t_out + o,
                             u_out + o * SBZ_BG_NB, du_out + o * SBZ_BG_NB, n_knots + i, n_rejected + i, t_fail + i);
}

#define BCHK(x)                                                                                   \
    do {                                                                                          \
        hipError_t e_ = (x);                                                                      \
        if (e_ != hipSuccess) {                                                                   \
            sbz_set_error("%s failed: %s", #x, hipGetErrorString(e_));                           \
            st = SBZ_ERR_HIP;                                                                     \
            goto done;                                                                            \
        }                                                                                         \
    } while (0)

extern "C" int sbz_background_solve_batch(double* params, int n, double rtol, double atol, int max_knots,
                                          double* t_out, double* u_out, double* du_out, int* n_knots,
                                          int* n_rejected, int* status, float* kernel_ms) {
    sbz_clear_error();
    if (!params || !t_out || !u_out || !du_out || !n_knots || !status || n < 1 || max_knots < 5 ||
        !(rtol > 0) || !(atol > 0)) {
        sbz_set_error("sbz_background_solve_batch: invalid argument");
        return SBZ_ERR_ARG;
    }
    int st = SBZ_OK;
    const size_t nk = (size_t)n * max_knots;
    double *dP = nullptr, *dt = nullptr, *du = nullptr, *ddu = nullptr, *dtf = nullptr;
    int *dn = nullptr, *dr = nullptr, *ds = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    std::vector<int> nrej(n);
    std::vector<double> tf(n);
    BCHK(hipMalloc(&dP, sizeof(double) * n * SBZ_NPAR));
    BCHK(hipMalloc(&dt, sizeof(double) * nk));
    BCHK(hipMalloc(&du, sizeof(double) * nk * SBZ_BG_NB));
    BCHK(hipMalloc(&ddu, sizeof(double) * nk * SBZ_BG_NB));
    BCHK(hipMalloc(&dtf, sizeof(double) * n));
    BCHK(hipMalloc(&dn, sizeof(int) * n));
    BCHK(hipMalloc(&dr, sizeof(int) * n));
    BCHK(hipMalloc(&ds, sizeof(int) * n));
    BCHK(hipMemcpy(dP, params, sizeof(double) * n * SBZ_NPAR, hipMemcpyHostToDevice));
    BCHK(hipEventCreate(&e0));
    BCHK(hipEventCreate(&e1));
    BCHK(hipEventRecord(e0, 0));
    hipLaunchKernelGGL(sbz_bg_solve_kernel, dim3((n + 63) / 64), dim3(64), 0, 0, dP, n, rtol, atol, max_knots, dt, du,
                       ddu, dn, dr, ds, dtf);
    BCHK(hipGetLastError());
    BCHK(hipEventRecord(e1, 0));
    BCHK(hipEventSynchronize(e1));
    if (kernel_ms) BCHK(hipEventElapsedTime(kernel_ms, e0, e1));
    BCHK(hipMemcpy(params, dP, sizeof(double) * n * SBZ_NPAR, hipMemcpyDeviceToHost));
    BCHK(hipMemcpy(n_knots, dn, sizeof(int) * n, hipMemcpyDeviceToHost));
    {
        // copy only the used prefix of every set's rows (entries past n_knots[i] are left as they are)
        int used = 1;
        for (int i = 0; i < n; ++i) used = n_knots[i] > used ? n_knots[i] : used;
        const size_t pt = sizeof(double) * max_knots, pu = pt * SBZ_BG_NB;
        BCHK(hipMemcpy2D(t_out, pt, dt, pt, sizeof(double) * used, n, hipMemcpyDeviceToHost));
        BCHK(hipMemcpy2D(u_out, pu, du, pu, sizeof(double) * used * SBZ_BG_NB, n, hipMemcpyDeviceToHost));
        BCHK(hipMemcpy2D(du_out, pu, ddu, pu, sizeof(double) * used * SBZ_BG_NB, n, hipMemcpyDeviceToHost));
    }
    BCHK(hipMemcpy(nrej.data(), dr, sizeof(int) * n, hipMemcpyDeviceToHost));
    BCHK(hipMemcpy(status, ds, sizeof(int) * n, hipMemcpyDeviceToHost));
    BCHK(hipMemcpy(tf.data(), dtf, sizeof(double) * n, hipMemcpyDeviceToHost));
    if (n_rejected)
        for (int i = 0; i < n; ++i) n_rejected[i] = nrej[i];
    for (int i = 0; i < n; ++i)
        if (status[i] != SBZ_BG_OK) {
            // per-set failures do not abort the batch; the first one is described
            sbz_set_error("background of set %d: %s at tau = %g", i, sbz_bg_status_text(status[i]), tf[i]);
            break;
        }
done:
    for (void* p : {(void*)dP, (void*)dt, (void*)du, (void*)ddu, (void*)dtf, (void*)dn, (void*)dr, (void*)ds})
        if (p) (void)hipFree(p);
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    return st;
}
